// a2c_table.h -- the residue templates of the full-atom protein builder, for
// device code (a2c.hip) and host code (a2c.hip's host part, and the host
// build in tests/a2c_table_host.cpp that checks them against the oracle).
//
// Each of the 20 amino acids is a chain of rigid groups hung on transforms
// M_child = M_parent * Ry(beta) * Tx(d) * Rx(alpha):
//   N  -- on the previous residue's C: alpha = omega of that residue,
//         beta = CA-C-N, d = C-N (the first residue's N is the origin, frame I);
//   CA -- on N: alpha = phi, beta = C-N-CA, d = N-CA (proline: its ring group);
//   side chain -- a linear chain of nodes from CA: a fixed "dummy" node, then
//         one node per side-chain group, alpha = xi1..xi5 or a constant;
//   C  -- on CA: alpha = psi, beta = N-CA-C, d = CA-C; the group (C, O).
// Atoms are numbered per residue N, CA (proline CA, CB, CG, CD), the side
// chain in node order, C, O.
//
// The numbers restate the reference's geometry and templates:
// Layers/FullAtomModel/cGeometry.cpp:10-93 (constants),
// Layers/FullAtomModel/cRigidGroup.cpp:29-203 (local atom positions) and
// Layers/FullAtomModel/cConformationAA.cpp:1-548 (nodes, angle binding, atom
// order).  The reference's carbonyl O and threonine OG1 sit at its serine C-O
// distance and isoleucine's CG2 at the full CG1-CB-CG2 angle; kept as there.
#pragma once

#include <stdint.h>

#ifdef __HIPCC__
#define A2C_HD __host__ __device__
#else
#define A2C_HD
#include <cmath>
#endif

namespace a2c {

constexpr double kPi = 3.14159265358979323846;

namespace geo {
constexpr double C_N_CA = kPi - 1.9391, N_CA_C = kPi - 2.061, CA_C_N = kPi - 2.1186;
constexpr double CA_C_O = kPi - 2.1033;
constexpr double R_CA_C = 1.525, R_C_N = 1.330, R_N_CA = 1.460;
constexpr double R_CA_CB = 1.52, C_CA_CB = -(kPi - kPi * 109.5 / 180.0);
constexpr double N_C_CA_CB = 0.5 * kPi * 122.6860 / 180.0, CORRECTION = 0.0;
constexpr double R_CB_OG = 1.417, CA_CB_OG = kPi * 110.773 / 180.0;
constexpr double R_CB_SG = 1.808, CA_CB_SG = kPi * 113.8169 / 180.0;
constexpr double R_CB_CG = 1.527, CA_CB_CG1 = kPi * 110.7 / 180.0, CG1_CB_CG2 = 1.88444687881;
constexpr double R_CG1_CD1 = 1.52, CB_CG1_CD1 = kPi * 113.97 / 180.0;
constexpr double CA_CB_OG1 = kPi * 109.18 / 180.0, OG1_CB_CG2 = 1.90291866134;
constexpr double R_C_C = R_CA_CB, C_C_C = C_CA_CB;
constexpr double R_CD_NE = 1.46, R_NE_CZ = 1.33, R_CZ_NH = 1.33;
constexpr double CG_CD_NE = -(kPi - kPi * 111.68 / 180.0);
constexpr double CD_NE_CZ = -(kPi - kPi * 124.79 / 180.0), NE_CZ_NH = kPi * 124.79 / 180.0;
constexpr double R_CD_CE = 1.46, R_CE_NZ = 1.33, CD_CE_NZ = -(kPi - kPi * 124.79 / 180.0);
constexpr double R_CG_OD = 1.25, CB_CG_OD = kPi * 119.22 / 180.0;
constexpr double R_CG_OD1 = 1.23, R_CG_ND2 = 1.33;
constexpr double CB_CG_OD1 = kPi * 120.85 / 180.0, CB_CG_ND2 = kPi * 116.48 / 180.0;
constexpr double R_CG_SD = 1.81, R_SD_CE = 1.79;
constexpr double CB_CG_SD = -(kPi - kPi * 112.69 / 180.0), CG_SD_CE = -(kPi - kPi * 100.61 / 180.0);
constexpr double R_CG_CD2 = 1.35, R_CG_ND1 = 1.38, R_CG_NE2 = 2.1912, R_CG_CE1 = 2.1915;
constexpr double CB_CG_CD2 = -(kPi - 2.27957453603), CB_CG_ND1 = kPi - 2.14413698608;
constexpr double CB_CG_NE2 = -(kPi - 2.90352974362), CB_CG_CE1 = kPi - 2.75209584734;
}  // namespace geo

constexpr int kTypes = 20;    // residue types, in the order of kCodes
constexpr int kMaxAtoms = 14; // atoms of the largest residue (tryptophan)
constexpr int kMaxSide = 6;   // side-chain nodes of the longest chain (arginine, lysine)
constexpr char kCodes[kTypes + 1] = "ACDEFGHIKLMNPQRSTVWY";

// How an atom's position in its group's frame follows from (r, a, b).
enum AtomKind : uint8_t {
    kOrigin = 0,  // (0, 0, 0)
    kLiteral,     // (r, a, b)
    kXZ,          // r (cos a, 0, sin a)
    kForkPlus,    // r (sin(a - pi/2),  cos(a - pi/2) sin b, cos(a - pi/2) cos b)
    kForkMinus,   // r (sin(a - pi/2), -cos(a - pi/2) sin b, cos(a - pi/2) cos b)
    kPlanePlus,   // r (sin(a - pi/2), 0,  cos(a - pi/2))
    kPlaneMinus,  // r (sin(a - pi/2), 0, -cos(a - pi/2))
    kRing         // r (cos a, sin a, 0)
};

struct Atom {
    uint8_t kind;
    double r, a, b;
    char name[4];
};

struct Side {
    int row;      // angle row bound to alpha (3..7 = xi1..xi5), -1: constant
    double alpha; // the constant alpha when row < 0
    double beta, d;
    int first, count;  // the node's atoms within the residue
};

struct Residue {
    char name3[4];
    int natoms, nca, nside;
};

#define A2C_O {kXZ, geo::R_CB_OG, geo::CA_C_O, 0, "O"}
#define A2C_Z(n) {kOrigin, 0, 0, 0, n}
#define A2C_L(n, x, y) {kLiteral, x, y, 0, n}
#define A2C_RING6 A2C_L("CD1", 0.69229424, -1.20533343), A2C_L("CE1", 2.08228309, -1.20846014), \
                  A2C_L("CZ", 2.77997771, -0.00625341), A2C_L("CE2", 2.08768347, 1.19908002), \
                  A2C_L("CD2", 0.69769462, 1.20220673)

// Residue names and counts, in the order of kCodes.
#define A2C_RESIDUES {                                                                           \
    {"ALA", 5, 1, 2}, {"CYS", 6, 1, 2}, {"ASP", 8, 1, 3}, {"GLU", 9, 1, 4}, {"PHE", 11, 1, 3},   \
    {"GLY", 4, 1, 0}, {"HIS", 10, 1, 3}, {"ILE", 8, 1, 3}, {"LYS", 9, 1, 6}, {"LEU", 8, 1, 3},   \
    {"MET", 8, 1, 5}, {"ASN", 8, 1, 3}, {"PRO", 7, 4, 0}, {"GLN", 9, 1, 4}, {"ARG", 11, 1, 6},   \
    {"SER", 6, 1, 2}, {"THR", 7, 1, 2}, {"VAL", 7, 1, 2}, {"TRP", 14, 1, 3}, {"TYR", 12, 1, 3}}

#define A2C_DUMMY {-1, geo::N_C_CA_CB, geo::CORRECTION, 0.0, 2, 0}
#define A2C_CB(count) {3, 0.0, geo::C_CA_CB, geo::R_CA_CB, 2, count}
#define A2C_CG {4, 0.0, geo::C_C_C, geo::R_C_C, 3, 1}
#define A2C_CD {5, 0.0, geo::C_C_C, geo::R_C_C, 4, 1}

// Side-chain nodes per residue (cConformationAA.cpp, add<Residue>).
#define A2C_SIDES {                                                                              \
    /* A */ {A2C_DUMMY, {-1, 0.0, geo::C_CA_CB, geo::R_CA_CB, 2, 1}},                            \
    /* C */ {A2C_DUMMY, A2C_CB(2)},                                                              \
    /* D */ {A2C_DUMMY, A2C_CB(1), {4, 0.0, geo::C_C_C, geo::R_C_C, 3, 3}},                      \
    /* E */ {A2C_DUMMY, A2C_CB(1), A2C_CG, {5, 0.0, geo::C_C_C, geo::R_C_C, 4, 3}},              \
    /* F */ {A2C_DUMMY, A2C_CB(1), {4, 0.0, geo::C_C_C, geo::R_C_C, 3, 6}},                      \
    /* G */ {},                                                                                  \
    /* H */ {A2C_DUMMY, A2C_CB(1), {4, 0.0, geo::C_C_C, geo::R_C_C, 3, 5}},                      \
    /* I */ {A2C_DUMMY, A2C_CB(2), {4, 0.0, geo::C_CA_CB, geo::R_CA_CB, 4, 2}},                  \
    /* K */ {A2C_DUMMY, A2C_CB(1), A2C_CG, A2C_CD, {6, 0.0, geo::C_C_C, geo::R_CD_CE, 5, 1},     \
             {7, 0.0, geo::CD_CE_NZ, geo::R_CE_NZ, 6, 1}},                                       \
    /* L */ {A2C_DUMMY, A2C_CB(1), {4, 0.0, geo::C_CA_CB, geo::R_CA_CB, 3, 3}},                  \
    /* M */ {A2C_DUMMY, A2C_CB(1), A2C_CG, {5, 0.0, geo::CB_CG_SD, geo::R_CG_SD, 4, 1},          \
             {6, 0.0, geo::CG_SD_CE, geo::R_SD_CE, 5, 1}},                                       \
    /* N */ {A2C_DUMMY, A2C_CB(1), {4, 0.0, geo::C_C_C, geo::R_C_C, 3, 3}},                      \
    /* P */ {},                                                                                  \
    /* Q */ {A2C_DUMMY, A2C_CB(1), A2C_CG, {5, 0.0, geo::C_C_C, geo::R_C_C, 4, 3}},              \
    /* R */ {A2C_DUMMY, A2C_CB(1), A2C_CG, A2C_CD, {6, 0.0, geo::CG_CD_NE, geo::R_CD_NE, 5, 1},  \
             {7, 0.0, geo::CD_NE_CZ, geo::R_NE_CZ, 6, 3}},                                       \
    /* S */ {A2C_DUMMY, A2C_CB(2)},                                                              \
    /* T */ {A2C_DUMMY, A2C_CB(3)},                                                              \
    /* V */ {A2C_DUMMY, A2C_CB(3)},                                                              \
    /* W */ {A2C_DUMMY, A2C_CB(1), {4, 0.0, geo::C_C_C, geo::R_C_C, 3, 9}},                      \
    /* Y */ {A2C_DUMMY, A2C_CB(1), {4, 0.0, geo::C_C_C, geo::R_C_C, 3, 7}}}

// Atoms per residue, in output order (cRigidGroup.cpp make<Group>).
#define A2C_ATOMS {                                                                              \
    /* A */ {A2C_Z("N"), A2C_Z("CA"), A2C_Z("CB"), A2C_Z("C"), A2C_O},                            \
    /* C */ {A2C_Z("N"), A2C_Z("CA"), A2C_Z("CB"),                                               \
             {kXZ, geo::R_CB_SG, -(kPi - geo::CA_CB_SG), 0, "SG"}, A2C_Z("C"), A2C_O},          \
    /* D */ {A2C_Z("N"), A2C_Z("CA"), A2C_Z("CB"), A2C_Z("CG"),                                  \
             {kPlanePlus, geo::R_CG_OD, geo::CB_CG_OD, 0, "OD1"},                               \
             {kPlaneMinus, geo::R_CG_OD, geo::CB_CG_OD, 0, "OD2"}, A2C_Z("C"), A2C_O},          \
    /* E */ {A2C_Z("N"), A2C_Z("CA"), A2C_Z("CB"), A2C_Z("CG"), A2C_Z("CD"),                     \
             {kPlanePlus, geo::R_CG_OD, geo::CB_CG_OD, 0, "OE1"},                               \
             {kPlaneMinus, geo::R_CG_OD, geo::CB_CG_OD, 0, "OE2"}, A2C_Z("C"), A2C_O},          \
    /* F */ {A2C_Z("N"), A2C_Z("CA"), A2C_Z("CB"), A2C_Z("CG"), A2C_RING6, A2C_Z("C"), A2C_O},   \
    /* G */ {A2C_Z("N"), A2C_Z("CA"), A2C_Z("C"), A2C_O},                                        \
    /* H */ {A2C_Z("N"), A2C_Z("CA"), A2C_Z("CB"), A2C_Z("CG"),                                  \
             {kRing, geo::R_CG_CD2, geo::CB_CG_CD2, 0, "CD2"},                                  \
             {kRing, geo::R_CG_NE2, geo::CB_CG_NE2, 0, "NE2"},                                  \
             {kRing, geo::R_CG_CE1, geo::CB_CG_CE1, 0, "CE1"},                                  \
             {kRing, geo::R_CG_ND1, geo::CB_CG_ND1, 0, "ND1"}, A2C_Z("C"), A2C_O},              \
    /* I */ {A2C_Z("N"), A2C_Z("CA"), A2C_Z("CB"),                                               \
             {kForkPlus, geo::R_CB_CG, geo::CA_CB_CG1, geo::CG1_CB_CG2, "CG2"}, A2C_Z("CG1"),   \
             {kXZ, geo::R_CG1_CD1, -(kPi - geo::CB_CG1_CD1), 0, "CD1"}, A2C_Z("C"), A2C_O},     \
    /* K */ {A2C_Z("N"), A2C_Z("CA"), A2C_Z("CB"), A2C_Z("CG"), A2C_Z("CD"), A2C_Z("CE"),        \
             A2C_Z("NZ"), A2C_Z("C"), A2C_O},                                                    \
    /* L */ {A2C_Z("N"), A2C_Z("CA"), A2C_Z("CB"), A2C_Z("CG"),                                  \
             {kForkPlus, geo::R_CB_CG, geo::CA_CB_CG1, geo::CG1_CB_CG2 / 2.0, "CD1"},           \
             {kForkMinus, geo::R_CB_CG, geo::CA_CB_CG1, geo::CG1_CB_CG2 / 2.0, "CD2"},          \
             A2C_Z("C"), A2C_O},                                                                 \
    /* M */ {A2C_Z("N"), A2C_Z("CA"), A2C_Z("CB"), A2C_Z("CG"), A2C_Z("SD"), A2C_Z("CE"),        \
             A2C_Z("C"), A2C_O},                                                                 \
    /* N */ {A2C_Z("N"), A2C_Z("CA"), A2C_Z("CB"), A2C_Z("CG"),                                  \
             {kPlanePlus, geo::R_CG_OD1, geo::CB_CG_OD1, 0, "OD1"},                             \
             {kPlaneMinus, geo::R_CG_ND2, geo::CB_CG_ND2, 0, "ND2"}, A2C_Z("C"), A2C_O},        \
    /* P */ {A2C_Z("N"), A2C_Z("CA"), {kLiteral, 0.36842372, -1.29540981, 0.70468246, "CB"},     \
             {kLiteral, -0.76137817, -2.21546205, 0.39299099, "CG"},                            \
             {kLiteral, -1.97838492, -1.33927458, 0.35849352, "CD"}, A2C_Z("C"), A2C_O},        \
    /* Q */ {A2C_Z("N"), A2C_Z("CA"), A2C_Z("CB"), A2C_Z("CG"), A2C_Z("CD"),                     \
             {kPlanePlus, geo::R_CG_OD1, geo::CB_CG_OD1, 0, "OE1"},                             \
             {kPlaneMinus, geo::R_CG_ND2, geo::CB_CG_ND2, 0, "NE2"}, A2C_Z("C"), A2C_O},        \
    /* R */ {A2C_Z("N"), A2C_Z("CA"), A2C_Z("CB"), A2C_Z("CG"), A2C_Z("CD"), A2C_Z("NE"),        \
             A2C_Z("CZ"), {kPlanePlus, geo::R_CZ_NH, geo::NE_CZ_NH, 0, "NH1"},                   \
             {kPlaneMinus, geo::R_CZ_NH, geo::NE_CZ_NH, 0, "NH2"}, A2C_Z("C"), A2C_O},          \
    /* S */ {A2C_Z("N"), A2C_Z("CA"), A2C_Z("CB"),                                               \
             {kXZ, geo::R_CB_OG, -(kPi - geo::CA_CB_OG), 0, "OG"}, A2C_Z("C"), A2C_O},          \
    /* T */ {A2C_Z("N"), A2C_Z("CA"), A2C_Z("CB"),                                               \
             {kForkPlus, geo::R_CB_OG, geo::CA_CB_OG1, geo::OG1_CB_CG2 / 2.0, "OG1"},           \
             {kForkMinus, geo::R_CB_CG, geo::CA_CB_CG1, geo::OG1_CB_CG2 / 2.0, "CG2"},          \
             A2C_Z("C"), A2C_O},                                                                 \
    /* V */ {A2C_Z("N"), A2C_Z("CA"), A2C_Z("CB"),                                               \
             {kForkPlus, geo::R_CB_CG, geo::CA_CB_CG1, geo::CG1_CB_CG2 / 2.0, "CG1"},           \
             {kForkMinus, geo::R_CB_CG, geo::CA_CB_CG1, geo::CG1_CB_CG2 / 2.0, "CG2"},          \
             A2C_Z("C"), A2C_O},                                                                 \
    /* W */ {A2C_Z("N"), A2C_Z("CA"), A2C_Z("CB"), A2C_Z("CG"),                                  \
             A2C_L("CD1", 0.829890578, -1.09003744), A2C_L("CD2", 0.84949676, 1.15028973),      \
             A2C_L("NE1", 2.1363699, -0.64570528), A2C_L("CE2", 2.18136443, 0.71889842),        \
             A2C_L("CE3", 0.61303306, 2.5301684), A2C_L("CZ2", 3.22086556, 1.65663275),         \
             A2C_L("CZ3", 1.68979906, 3.42486545), A2C_L("CH2", 2.92849902, 3.0257584),         \
             A2C_Z("C"), A2C_O},                                                                 \
    /* Y */ {A2C_Z("N"), A2C_Z("CA"), A2C_Z("CB"), A2C_Z("CG"), A2C_RING6,                       \
             A2C_L("OH", 4.16960749, 0.05377697), A2C_Z("C"), A2C_O}}

// Residue type of a one-letter code, -1 outside the 20 amino acids.
A2C_HD inline int residue_type(uint8_t c) {
    switch (c) {
        case 'A': return 0;  case 'C': return 1;  case 'D': return 2;  case 'E': return 3;
        case 'F': return 4;  case 'G': return 5;  case 'H': return 6;  case 'I': return 7;
        case 'K': return 8;  case 'L': return 9;  case 'M': return 10; case 'N': return 11;
        case 'P': return 12; case 'Q': return 13; case 'R': return 14; case 'S': return 15;
        case 'T': return 16; case 'V': return 17; case 'W': return 18; case 'Y': return 19;
        default: return -1;
    }
}

// An atom's position in its group's frame, with sc(x, &sin, &cos) giving the
// sine and cosine.
template <typename SinCos>
A2C_HD inline void local_position_with(const Atom& at, double v[3], SinCos sc) {
    const double h = at.a - kPi / 2.0;
    double sa, ca, sh, ch, sb, cb;
    switch (at.kind) {
        case kLiteral: v[0] = at.r; v[1] = at.a; v[2] = at.b; return;
        case kXZ: sc(at.a, &sa, &ca); v[0] = at.r * ca; v[1] = 0.0; v[2] = at.r * sa; return;
        case kForkPlus:
        case kForkMinus: {
            const double s = at.kind == kForkPlus ? 1.0 : -1.0;
            sc(h, &sh, &ch);
            sc(at.b, &sb, &cb);
            v[0] = at.r * sh;
            v[1] = s * at.r * ch * sb;
            v[2] = at.r * ch * cb;
            return;
        }
        case kPlanePlus:
        case kPlaneMinus:
            sc(h, &sh, &ch);
            v[0] = at.r * sh;
            v[1] = 0.0;
            v[2] = (at.kind == kPlanePlus ? 1.0 : -1.0) * at.r * ch;
            return;
        case kRing: sc(at.a, &sa, &ca); v[0] = at.r * ca; v[1] = at.r * sa; v[2] = 0.0; return;
        default: v[0] = v[1] = v[2] = 0.0; return;
    }
}

struct LibSinCos {
    A2C_HD void operator()(double x, double* s, double* c) const {
        *s = sin(x);
        *c = cos(x);
    }
};

// An atom's position in its group's frame (the math library's sin and cos).
A2C_HD inline void local_position(const Atom& at, double v[3]) {
    local_position_with(at, v, LibSinCos{});
}

}  // namespace a2c
