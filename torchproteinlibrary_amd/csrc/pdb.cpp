// pdb.cpp -- the PDB reader behind include/tpl_pdb.h (host code).
//
// Fixed-column parsing of ATOM records as the reference's cPDBLoader does
// (Layers/FullAtomModel/cPDBLoader.cpp:23-53): heavy atoms only (name starting
// with C, N, O, S; Math/nUtil.cpp:166-171), coordinates parsed as float and
// widened (std::stof there), names trimmed.  The file is read once into
// memory and scanned line by line.
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "tpl_pdb.h"

namespace {

struct Field {
    const char* p;
    int n;
};

// Columns [a, a + n) of a line of length len, clipped to the line.
Field cols(const char* line, int len, int a, int n) {
    if (a >= len) return {line + len, 0};
    return {line + a, n < len - a ? n : len - a};
}

Field trim(Field f) {
    while (f.n > 0 && std::isspace((unsigned char)f.p[0])) ++f.p, --f.n;
    while (f.n > 0 && std::isspace((unsigned char)f.p[f.n - 1])) --f.n;
    return f;
}

void put_name(uint8_t* dst, Field f) {
    std::memset(dst, 0, 4);
    std::memcpy(dst, f.p, f.n < 4 ? f.n : 4);
}

bool parse_float(Field f, float* out) {
    char buf[32];
    if (f.n <= 0 || f.n >= (int)sizeof(buf)) return false;
    std::memcpy(buf, f.p, f.n);
    buf[f.n] = 0;
    char* end = nullptr;
    *out = std::strtof(buf, &end);
    return end != buf;
}

bool parse_int(Field f, int32_t* out) {
    char buf[32];
    if (f.n <= 0 || f.n >= (int)sizeof(buf)) return false;
    std::memcpy(buf, f.p, f.n);
    buf[f.n] = 0;
    char* end = nullptr;
    const long v = std::strtol(buf, &end, 10);
    if (end == buf) return false;
    *out = (int32_t)v;
    return true;
}

bool read_file(const char* filename, std::string& text) {
    FILE* f = std::fopen(filename, "rb");
    if (!f) return false;
    char buf[1 << 16];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, n);
    std::fclose(f);
    return true;
}

// Calls rec(line, len) for every heavy-atom ATOM record; stops early when rec
// returns a nonzero code, which it returns.
template <typename F>
int scan(const std::string& text, F rec) {
    const char* p = text.data();
    const char* end = p + text.size();
    while (p < end) {
        const char* nl = static_cast<const char*>(std::memchr(p, '\n', end - p));
        const char* le = nl ? nl : end;
        int len = (int)(le - p);
        if (len > 0 && p[len - 1] == '\r') --len;
        if (len >= 4 && std::memcmp(p, "ATOM", 4) == 0) {
            const Field name = trim(cols(p, len, 12, 4));
            if (name.n > 0 && std::strchr("CNOS", name.p[0])) {
                const int rc = rec(p, len, name);
                if (rc) return rc;
            }
        }
        p = nl ? nl + 1 : end;
    }
    return 0;
}

}  // namespace

extern "C" {

const char* tpl_pdb_strerror(int code) {
    switch (code) {
        case TPL_PDB_ERR_OPEN: return "cannot open the file";
        case TPL_PDB_ERR_RECORD: return "a heavy-atom ATOM record is too short or unparsable";
        case TPL_PDB_ERR_ARG: return "invalid argument (null pointer or negative size)";
        default: return code >= 0 ? "ok" : "unknown error code";
    }
}

int tpl_pdb_count(const char* filename) {
    if (!filename) return TPL_PDB_ERR_ARG;
    std::string text;
    if (!read_file(filename, text)) return TPL_PDB_ERR_OPEN;
    int n = 0;
    scan(text, [&](const char*, int, Field) {
        ++n;
        return 0;
    });
    return n;
}

int tpl_pdb_read(const char* filename, double* coords, uint8_t* chain_names, uint8_t* res_names,
                 int32_t* res_nums, uint8_t* atom_names, int max_atoms) {
    if (!filename || max_atoms < 0) return TPL_PDB_ERR_ARG;
    if (max_atoms > 0 && (!coords || !chain_names || !res_names || !res_nums || !atom_names))
        return TPL_PDB_ERR_ARG;
    std::string text;
    if (!read_file(filename, text)) return TPL_PDB_ERR_OPEN;
    int n = 0;
    const int rc = scan(text, [&](const char* line, int len, Field name) {
        if (n >= max_atoms) return 1;
        float x, y, z;
        int32_t num;
        if (!parse_float(cols(line, len, 30, 8), &x) || !parse_float(cols(line, len, 38, 8), &y) ||
            !parse_float(cols(line, len, 46, 8), &z) || !parse_int(cols(line, len, 22, 4), &num))
            return (int)TPL_PDB_ERR_RECORD;
        coords[3 * n] = x;
        coords[3 * n + 1] = y;
        coords[3 * n + 2] = z;
        put_name(atom_names + 4 * n, name);
        put_name(res_names + 4 * n, trim(cols(line, len, 17, 3)));
        put_name(chain_names + 4 * n, trim(cols(line, len, 21, 1)));
        res_nums[n] = num;
        ++n;
        return 0;
    });
    return rc < 0 ? rc : n;
}

}  // extern "C"
