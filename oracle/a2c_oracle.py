"""FP64 CPU restatement of TorchProteinLibrary's FullAtomModel.Angles2Coords.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg as the checker, never by the product path.

What it restates:
* the geometry constants (Layers/FullAtomModel/cGeometry.cpp:10-93);
* the rigid groups and their local atom positions
  (Layers/FullAtomModel/cRigidGroup.cpp:29-236);
* the per-residue trees of transforms, atom order and angle binding
  (Layers/FullAtomModel/cConformationAA.cpp:1-548, cConformation.cpp:29-125):
  a node's frame is M = M_parent * Ry(beta) * Tx(d) * Rx(alpha)
  (cConformation.cpp:7-13, 142-158), the root's frame is the identity;
* the backward (cConformation.cpp:160-195): for every node whose alpha is an
  input angle, dL/dalpha = sum over the atoms x of the node's subtree of
  g . (F x) with F = M_parent * dTransform * M^-1, restated here with the 4x4
  matrices as written there.

The reference keeps the quirks restated below (the carbonyl O and threonine
OG1 at the serine C-O distance, isoleucine's CG2 at the full CG1-CB-CG2
angle); they are part of its output and are kept.

Parity of this module is pinned against the reference's own compiled CPU code
(oracle/_ref/libtplref_a2c.so) and against geometry and finite differences in
tests/test_a2c_oracle.py.
"""
import numpy as np

PI = np.pi

# cGeometry.cpp:10-93 (names without the _angle suffix)
GEO = {}
GEO["C_N_CA"] = PI - 1.9391
GEO["N_CA_C"] = PI - 2.061
GEO["CA_C_N"] = PI - 2.1186
GEO["CA_C_O"] = PI - 2.1033
GEO["R_CA_C"], GEO["R_C_N"], GEO["R_N_CA"] = 1.525, 1.330, 1.460
GEO["R_CA_CB"] = 1.52
GEO["C_CA_CB"] = -(PI - PI * 109.5 / 180.0)
GEO["N_C_CA_CB"] = 0.5 * PI * 122.6860 / 180.0
GEO["correction"] = 0.0
GEO["R_CB_OG"], GEO["CA_CB_OG"] = 1.417, PI * 110.773 / 180.0
GEO["R_CB_SG"], GEO["CA_CB_SG"] = 1.808, PI * 113.8169 / 180.0
GEO["R_CB_CG"], GEO["CA_CB_CG1"], GEO["CG1_CB_CG2"] = 1.527, PI * 110.7 / 180.0, 1.88444687881
GEO["R_CG1_CD1"], GEO["CB_CG1_CD1"] = 1.52, PI * 113.97 / 180.0
GEO["CA_CB_OG1"], GEO["OG1_CB_CG2"] = PI * 109.18 / 180.0, 1.90291866134
GEO["R_C_C"], GEO["C_C_C"] = GEO["R_CA_CB"], GEO["C_CA_CB"]
GEO["R_CD_NE"], GEO["R_NE_CZ"], GEO["R_CZ_NH"] = 1.46, 1.33, 1.33
GEO["CG_CD_NE"] = -(PI - PI * 111.68 / 180.0)
GEO["CD_NE_CZ"] = -(PI - PI * 124.79 / 180.0)
GEO["NE_CZ_NH"] = PI * 124.79 / 180.0
GEO["R_CD_CE"], GEO["R_CE_NZ"] = 1.46, 1.33
GEO["CD_CE_NZ"] = -(PI - PI * 124.79 / 180.0)
GEO["R_CG_OD"], GEO["CB_CG_OD"] = 1.25, PI * 119.22 / 180.0
GEO["R_CG_OD1"], GEO["R_CG_ND2"] = 1.23, 1.33
GEO["CB_CG_OD1"], GEO["CB_CG_ND2"] = PI * 120.85 / 180.0, PI * 116.48 / 180.0
GEO["R_CG_SD"], GEO["R_SD_CE"] = 1.81, 1.79
GEO["CB_CG_SD"] = -(PI - PI * 112.69 / 180.0)
GEO["CG_SD_CE"] = -(PI - PI * 100.61 / 180.0)
GEO["R_CG_CD2"], GEO["R_CG_ND1"], GEO["R_CG_NE2"], GEO["R_CG_CE1"] = 1.35, 1.38, 2.1912, 2.1915
GEO["CB_CG_CD2"] = -(PI - 2.27957453603)
GEO["CB_CG_ND1"] = PI - 2.14413698608
GEO["CB_CG_NE2"] = -(PI - 2.90352974362)
GEO["CB_CG_CE1"] = PI - 2.75209584734


def _xz(r, a):  # r * (cos a, 0, sin a)
    return np.array([r * np.cos(a), 0.0, r * np.sin(a)])


def _fork(r, a, b, s):  # r * (sin(a - pi/2), s cos(a - pi/2) sin b, cos(a - pi/2) cos b)
    h = a - PI / 2.0
    return r * np.array([np.sin(h), s * np.cos(h) * np.sin(b), np.cos(h) * np.cos(b)])


def _plane(r, a, s):  # r * (sin(a - pi/2), 0, s cos(a - pi/2))
    h = a - PI / 2.0
    return r * np.array([np.sin(h), 0.0, s * np.cos(h)])


def _ring(r, a):  # r * (cos a, sin a, 0)
    return r * np.array([np.cos(a), np.sin(a), 0.0])


Z = np.zeros(3)
G = GEO
# Rigid groups (cRigidGroup.cpp): name -> [(atom name, local position)]
_VAL = [("CB", Z), ("CG1", _fork(G["R_CB_CG"], G["CA_CB_CG1"], G["CG1_CB_CG2"] / 2, 1)),
        ("CG2", _fork(G["R_CB_CG"], G["CA_CB_CG1"], G["CG1_CB_CG2"] / 2, -1))]
_RING6 = [(0.69229424, -1.20533343), (2.08228309, -1.20846014), (2.77997771, -0.00625341),
          (2.08768347, 1.19908002), (0.69769462, 1.20220673)]


def _lit(names, xy):
    return [(n, np.array([x, y, 0.0])) for n, (x, y) in zip(names, xy)]


GROUPS = {
    "carbonyl": [("C", Z), ("O", _xz(G["R_CB_OG"], G["CA_C_O"]))],  # :33-46 (R_CB_OG as there)
    "ser": [("CB", Z), ("OG", _xz(G["R_CB_OG"], -(PI - G["CA_CB_OG"])))],  # :48-53
    "cys": [("CB", Z), ("SG", _xz(G["R_CB_SG"], -(PI - G["CA_CB_SG"])))],  # :55-60
    "val": _VAL,  # :62-72
    "ile1": [("CB", Z), ("CG2", _fork(G["R_CB_CG"], G["CA_CB_CG1"], G["CG1_CB_CG2"], 1))],  # :74-80
    "ile2": [("CG1", Z), ("CD1", _xz(G["R_CG1_CD1"], -(PI - G["CB_CG1_CD1"])))],  # :82-87
    "leu": [("CG", Z), ("CD1", _VAL[1][1]), ("CD2", _VAL[2][1])],  # :89-99
    "thr": [("CB", Z), ("OG1", _fork(G["R_CB_OG"], G["CA_CB_OG1"], G["OG1_CB_CG2"] / 2, 1)),
            ("CG2", _fork(G["R_CB_CG"], G["CA_CB_CG1"], G["OG1_CB_CG2"] / 2, -1))],  # :101-111
    "arg": [("CZ", Z), ("NH1", _plane(G["R_CZ_NH"], G["NE_CZ_NH"], 1)),
            ("NH2", _plane(G["R_CZ_NH"], G["NE_CZ_NH"], -1))],  # :113-123
    "asp": [("CG", Z), ("OD1", _plane(G["R_CG_OD"], G["CB_CG_OD"], 1)),
            ("OD2", _plane(G["R_CG_OD"], G["CB_CG_OD"], -1))],  # :125-135
    "asn": [("CG", Z), ("OD1", _plane(G["R_CG_OD1"], G["CB_CG_OD1"], 1)),
            ("ND2", _plane(G["R_CG_ND2"], G["CB_CG_ND2"], -1))],  # :137-147
    "his": [("CG", Z), ("CD2", _ring(G["R_CG_CD2"], G["CB_CG_CD2"])),
            ("NE2", _ring(G["R_CG_NE2"], G["CB_CG_NE2"])), ("CE1", _ring(G["R_CG_CE1"], G["CB_CG_CE1"])),
            ("ND1", _ring(G["R_CG_ND1"], G["CB_CG_ND1"]))],  # :149-157
    "pro": [("CA", Z)] + [(n, np.array(v)) for n, v in zip(
        ("CB", "CG", "CD"), ((0.36842372, -1.29540981, 0.70468246),
                             (-0.76137817, -2.21546205, 0.39299099),
                             (-1.97838492, -1.33927458, 0.35849352)))],  # :159-166
    "phe": [("CG", Z)] + _lit(("CD1", "CE1", "CZ", "CE2", "CD2"), _RING6),  # :168-177
    "tyr": [("CG", Z)] + _lit(("CD1", "CE1", "CZ", "CE2", "CD2"), _RING6)
           + [("OH", np.array([4.16960749, 0.05377697, 0.0]))],  # :179-189
    "trp": [("CG", Z)] + _lit(("CD1", "CD2", "NE1", "CE2", "CE3", "CZ2", "CZ3", "CH2"), (
        (0.829890578, -1.09003744), (0.84949676, 1.15028973), (2.1363699, -0.64570528),
        (2.18136443, 0.71889842), (0.61303306, 2.5301684), (3.22086556, 1.65663275),
        (1.68979906, 3.42486545), (2.92849902, 3.0257584))),  # :191-203
}


def _atom(name):
    return [(name, Z)]


# Side chains (cConformationAA.cpp): list of (parent, alpha, beta, d, group)
# with parent "CA" or the index of an earlier side-chain node, alpha an angle
# row (int: 3..7 = xi1..xi5) or a constant (float).
_DUMMY = ("CA", G["N_C_CA_CB"], G["correction"], 0.0, [])  # ADD_DUMMY_TRANSFORM
_CB = lambda grp: (0, 3, G["C_CA_CB"], G["R_CA_CB"], grp)  # noqa: E731  bbCB on the dummy
_CG = (1, 4, G["C_C_C"], G["R_C_C"], _atom("CG"))  # ADD_CARBON_GAMMA
_CD = (2, 5, G["C_C_C"], G["R_C_C"], _atom("CD"))  # ADD_CARBON_DELTA
SIDE = {
    "G": [],
    "A": [_DUMMY, (0, 0.0, G["C_CA_CB"], G["R_CA_CB"], _atom("CB"))],
    "S": [_DUMMY, _CB(GROUPS["ser"])],
    "C": [_DUMMY, _CB(GROUPS["cys"])],
    "V": [_DUMMY, _CB(GROUPS["val"])],
    "I": [_DUMMY, _CB(GROUPS["ile1"]), (1, 4, G["C_CA_CB"], G["R_CA_CB"], GROUPS["ile2"])],
    "L": [_DUMMY, _CB(_atom("CB")), (1, 4, G["C_CA_CB"], G["R_CA_CB"], GROUPS["leu"])],
    "T": [_DUMMY, _CB(GROUPS["thr"])],
    "R": [_DUMMY, _CB(_atom("CB")), _CG, _CD, (3, 6, G["CG_CD_NE"], G["R_CD_NE"], _atom("NE")),
          (4, 7, G["CD_NE_CZ"], G["R_NE_CZ"], GROUPS["arg"])],
    "K": [_DUMMY, _CB(_atom("CB")), _CG, _CD, (3, 6, G["C_C_C"], G["R_CD_CE"], _atom("CE")),
          (4, 7, G["CD_CE_NZ"], G["R_CE_NZ"], _atom("NZ"))],
    "D": [_DUMMY, _CB(_atom("CB")), (1, 4, G["C_C_C"], G["R_C_C"], GROUPS["asp"])],
    "N": [_DUMMY, _CB(_atom("CB")), (1, 4, G["C_C_C"], G["R_C_C"], GROUPS["asn"])],
    "E": [_DUMMY, _CB(_atom("CB")), _CG,
          (2, 5, G["C_C_C"], G["R_C_C"], [(n.replace("G", "D").replace("OD", "OE"), v)
                                           for n, v in GROUPS["asp"]])],
    "Q": [_DUMMY, _CB(_atom("CB")), _CG,
          (2, 5, G["C_C_C"], G["R_C_C"], list(zip(("CD", "OE1", "NE2"), [v for _, v in GROUPS["asn"]])))],
    "M": [_DUMMY, _CB(_atom("CB")), _CG, (2, 5, G["CB_CG_SD"], G["R_CG_SD"], _atom("SD")),
          (3, 6, G["CG_SD_CE"], G["R_SD_CE"], _atom("CE"))],
    "H": [_DUMMY, _CB(_atom("CB")), (1, 4, G["C_C_C"], G["R_C_C"], GROUPS["his"])],
    "P": [],
    "F": [_DUMMY, _CB(_atom("CB")), (1, 4, G["C_C_C"], G["R_C_C"], GROUPS["phe"])],
    "Y": [_DUMMY, _CB(_atom("CB")), (1, 4, G["C_C_C"], G["R_C_C"], GROUPS["tyr"])],
    "W": [_DUMMY, _CB(_atom("CB")), (1, 4, G["C_C_C"], G["R_C_C"], GROUPS["trp"])],
}
RES3 = {"G": "GLY", "A": "ALA", "S": "SER", "C": "CYS", "V": "VAL", "I": "ILE", "L": "LEU",
        "T": "THR", "R": "ARG", "K": "LYS", "D": "ASP", "N": "ASN", "E": "GLU", "Q": "GLN",
        "M": "MET", "H": "HIS", "P": "PRO", "F": "PHE", "Y": "TYR", "W": "TRP"}  # nUtil.cpp:519-562


def residue_atoms(aa):
    """Atom names of one residue in output order: N, CA (proline: its ring
    group), side chain in node order, C, O (cConformationAA.cpp)."""
    ca = [n for n, _ in GROUPS["pro"]] if aa == "P" else ["CA"]
    sc = [n for node in SIDE[aa] for n, _ in node[4]]
    return ["N"] + ca + sc + ["C", "O"]


def num_atoms(seq):
    """Atoms of a sequence (nUtil.cpp:174-190 with add_terminal false)."""
    return sum(len(residue_atoms(a)) for a in seq)


def _rx(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[1, 0, 0, 0], [0, c, -s, 0], [0, s, c, 0], [0, 0, 0, 1.0]])


def _drx(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[0, 0, 0, 0], [0, -s, -c, 0], [0, c, -s, 0], [0, 0, 0, 0.0]])


def _ry(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, 0, s, 0], [0, 1, 0, 0], [-s, 0, c, 0], [0, 0, 0, 1.0]])


def _tx(d):
    m = np.eye(4)
    m[0, 3] = d
    return m


def _inv(m):  # rigid inverse (cMatrix44.cpp:142-153)
    r = np.eye(4)
    r[:3, :3] = m[:3, :3].T
    r[:3, 3] = -m[:3, :3].T @ m[:3, 3]
    return r


def build(seq):
    """The conformation tree of a sequence (cConformation.cpp:29-125):
    nodes as dicts with parent (index or -1), alpha source ((row, col) of the
    angle array or a constant), beta, d, local atoms and their output indices."""
    nodes = []
    last_c = -1
    atom = 0

    def add(parent, alpha, beta, d, group):
        nonlocal atom
        idx = list(range(atom, atom + len(group)))
        atom += len(group)
        nodes.append(dict(parent=parent, alpha=alpha, beta=beta, d=d,
                          local=[v for _, v in group], names=[n for n, _ in group], index=idx))
        return len(nodes) - 1

    for i, aa in enumerate(seq):
        if aa not in SIDE:
            raise ValueError(f"unknown residue letter {aa!r}")
        if last_c < 0:  # ADD_NITROGEN at the root: identity transform
            n_node = add(-1, 0.0, 0.0, 0.0, _atom("N"))
        else:  # omega of the previous residue
            n_node = add(last_c, (2, i - 1), G["CA_C_N"], G["R_C_N"], _atom("N"))
        ca_group = GROUPS["pro"] if aa == "P" else _atom("CA")
        ca = add(n_node, (0, i), G["C_N_CA"], G["R_N_CA"], ca_group)
        sc_ids = []
        for parent, alpha, beta, d, group in SIDE[aa]:
            par = ca if parent == "CA" else sc_ids[parent]
            al = (alpha, i) if isinstance(alpha, int) else alpha
            sc_ids.append(add(par, al, beta, d, group))
        last_c = add(ca, (1, i), G["N_CA_C"], G["R_CA_C"], GROUPS["carbonyl"])
    return nodes, atom


def _alpha(node, angles):
    a = node["alpha"]
    return angles[a[0], a[1]] if isinstance(a, tuple) else a


def forward_one(seq, angles):
    """coords [3 * num_atoms] and the node frames for angles [8, >= len(seq)]
    (cConformation.cpp:142-158)."""
    nodes, n = build(seq)
    coords = np.zeros(3 * n)
    frames = []
    for node in nodes:
        if node["parent"] < 0:
            m = np.eye(4)
        else:
            m = frames[node["parent"]] @ _ry(node["beta"]) @ _tx(node["d"]) @ _rx(_alpha(node, angles))
        frames.append(m)
        for k, loc in zip(node["index"], node["local"]):
            coords[3 * k:3 * k + 3] = m[:3, :3] @ loc + m[:3, 3]
    return coords, nodes, frames


def backward_one(seq, angles, grad_coords):
    """grad_angles [8, stride] (cConformation.cpp:160-195): for a node bound to
    an input angle, sum over its subtree's atoms of g . (F x)."""
    coords, nodes, frames = forward_one(seq, angles)
    children = [[] for _ in nodes]
    for j, node in enumerate(nodes):
        if node["parent"] >= 0:
            children[node["parent"]].append(j)
    grad = np.zeros_like(angles, dtype=np.float64)
    x = coords.reshape(-1, 3)
    g = np.asarray(grad_coords[:coords.size], dtype=np.float64).reshape(-1, 3)
    for j, node in enumerate(nodes):
        if not isinstance(node["alpha"], tuple):
            continue
        dmat = _ry(node["beta"]) @ _tx(node["d"]) @ _drx(_alpha(node, angles))
        F = frames[node["parent"]] @ dmat @ _inv(frames[j])
        sub, stack = [], [j]
        while stack:
            k = stack.pop()
            sub += nodes[k]["index"]
            stack += children[k]
        xs = x[sub]
        fx = xs @ F[:3, :3].T + F[:3, 3]
        grad[node["alpha"]] = np.sum(g[sub] * fx)
    return grad


def names_one(seq):
    """(residue names, atom names) per atom, as strings."""
    res, atoms = [], []
    for aa in seq:
        a = residue_atoms(aa)
        atoms += a
        res += [RES3[aa]] * len(a)
    return res, atoms


def forward(sequences, angles):
    """Batch: angles [B, 8, L] -> coords [B, 3 * max atoms] (0 past each
    structure's atoms), num_atoms [B]."""
    counts = [num_atoms(s) for s in sequences]
    out = np.zeros((len(sequences), 3 * max(counts + [0])))
    for b, s in enumerate(sequences):
        c, _, _ = forward_one(s, angles[b])
        out[b, :c.size] = c
    return out, np.array(counts, dtype=np.int32)


def backward(sequences, angles, grad_coords):
    grad = np.zeros(angles.shape)
    for b, s in enumerate(sequences):
        grad[b] = backward_one(s, angles[b], grad_coords[b])
    return grad


AA20 = "ACDEFGHIKLMNPQRSTVWY"


def workload(B, seed, min_len=30, max_len=1000, dtype=np.float64):
    """Random sequences with log-uniform lengths and angles [B, 8, Lmax]:
    phi ~ N(-1.1, 0.3), psi ~ N(2.3, 0.4), omega = pi + N(0, 0.05), side-chain
    angles ~ U(-pi, pi); 0 past each sequence."""
    rng = np.random.default_rng(seed)
    lens = np.floor(np.exp(rng.uniform(np.log(min_len), np.log(max_len + 1), size=B))).astype(int)
    lens = np.clip(lens, min_len, max_len)
    seqs = ["".join(rng.choice(list(AA20), size=n)) for n in lens]
    L = int(lens.max())
    ang = np.zeros((B, 8, L))
    for b, n in enumerate(lens):
        ang[b, 0, :n] = rng.normal(-1.1, 0.3, n)
        ang[b, 1, :n] = rng.normal(2.3, 0.4, n)
        ang[b, 2, :n] = PI + rng.normal(0, 0.05, n)
        ang[b, 3:, :n] = rng.uniform(-PI, PI, (5, n))
    return seqs, ang.astype(dtype).astype(np.float64)


def backward_fast(seq, angles, grad_coords):
    """The same gradient in O(atoms): subtree sums (sum g, sum x cross g)
    accumulated from the leaves (every node comes after its parent in build
    order), then dL/dalpha = u . (sum x cross g - p cross sum g) with u the
    node frame's x axis and p its origin (F x = u cross (x - p) for the
    reference's F = M_parent dT M^-1, cConformation.cpp:148).  Pinned to
    backward_one in tests/test_a2c_oracle.py; used for long chains."""
    coords, nodes, frames = forward_one(seq, angles)
    x = coords.reshape(-1, 3)
    g = np.asarray(grad_coords[:coords.size], dtype=np.float64).reshape(-1, 3)
    S1 = np.zeros((len(nodes), 3))
    S2 = np.zeros((len(nodes), 3))
    for j, node in enumerate(nodes):
        idx = node["index"]
        if idx:
            S1[j] = g[idx].sum(axis=0)
            S2[j] = np.cross(x[idx], g[idx]).sum(axis=0)
    for j in range(len(nodes) - 1, 0, -1):
        par = nodes[j]["parent"]
        if par >= 0:
            S1[par] += S1[j]
            S2[par] += S2[j]
    grad = np.zeros_like(angles, dtype=np.float64)
    for j, node in enumerate(nodes):
        if isinstance(node["alpha"], tuple):
            u, p = frames[j][:3, 0], frames[j][:3, 3]
            grad[node["alpha"]] = u @ (S2[j] - np.cross(p, S1[j]))
    return grad
