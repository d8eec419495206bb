// SDF ("self-describing format") snapshots of the body table: the file format
// 2HOT reads and writes its particles in (nln/output.c:270-423 writes the
// outbody table through libSDF's SDFwrite64; libSDF itself is not part of
// the reference tree).
//
// Layout, as the reference's own SDF writer in sph+nln/wind.c:82-135 emits it:
//
//   # SDF
//   parameter byteorder = 0x78563412;      (little-endian)
//   int name = %d;  int64_t name = %lld;  float name = %g;  double name = %.16g;
//   char name[] = "%s";                    (one declaration per header value)
//   struct { float mass; float x, y, z; ... }[N];
//   #\f
//   # SDF-EOH <spaces up to a 32-byte boundary>\n
//   N packed records
//
// The reader parses that grammar: scalar declarations (int, int64_t, float,
// double, char[]), C comments, and one struct declaration whose members give
// the record layout (packed; char 1, short 2, int / float 4, int64_t / long /
// double 8 bytes).  Big-endian files are refused.
#include <ctype.h>
#include <stdio.h>
#include <string.h>

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "pgrav.h"

namespace pgrav {

void fail(const char *fmt, ...);  // error reporting of the library (pgrav_abi.hip)
void clear_error();

namespace {

const int DATAALIGN = 32;  // sph+nln/wind.c:36

struct Value {
    std::string name, s;
    int type = PGRAV_SDF_INT;
    int64_t i = 0;
    double d = 0;
};

struct Field {
    std::string name;
    int type;
    int64_t offset;
};

struct SdfFile {
    std::string path;
    std::vector<Value> values;
    std::vector<Field> fields;
    std::string desc;
    int64_t recsize = 0, nrec = -1, data_offset = 0;
};

int type_size(int t) {
    switch (t) {
    case PGRAV_SDF_CHAR: return 1;
    case PGRAV_SDF_SHORT: return 2;
    case PGRAV_SDF_INT:
    case PGRAV_SDF_FLOAT: return 4;
    default: return 8;  // int64, double
    }
}

bool type_of(const std::string &w, int *t) {
    static const struct {
        const char *w;
        int t;
    } tab[] = {{"char", PGRAV_SDF_CHAR},     {"short", PGRAV_SDF_SHORT},  {"int", PGRAV_SDF_INT},
               {"int32_t", PGRAV_SDF_INT},   {"float", PGRAV_SDF_FLOAT},  {"double", PGRAV_SDF_DOUBLE},
               {"int64_t", PGRAV_SDF_INT64}, {"long", PGRAV_SDF_INT64},   {"int64", PGRAV_SDF_INT64}};
    for (auto &e : tab)
        if (w == e.w) {
            *t = e.t;
            return true;
        }
    return false;
}

// tokens of the header: identifiers, numbers, strings and punctuation
struct Lexer {
    const std::string &s;
    size_t p = 0;
    explicit Lexer(const std::string &src) : s(src) {}
    void skip() {
        for (;;) {
            while (p < s.size() && isspace((unsigned char)s[p])) p++;
            if (p + 1 < s.size() && s[p] == '/' && s[p + 1] == '*') {
                size_t e = s.find("*/", p + 2);
                p = e == std::string::npos ? s.size() : e + 2;
            } else if (p < s.size() && s[p] == '#') {  // comment line
                while (p < s.size() && s[p] != '\n') p++;
            } else {
                return;
            }
        }
    }
    std::string next() {
        skip();
        if (p >= s.size())
            return "";
        const size_t b = p;
        if (s[p] == '"') {
            size_t e = s.find('"', p + 1);
            if (e == std::string::npos)
                throw std::runtime_error("SDF header: unterminated string");
            p = e + 1;
            return s.substr(b, p - b);
        }
        if (isalnum((unsigned char)s[p]) || s[p] == '_' || s[p] == '-' || s[p] == '+' || s[p] == '.') {
            while (p < s.size() && (isalnum((unsigned char)s[p]) || strchr("_-+.", s[p]))) p++;
            return s.substr(b, p - b);
        }
        p++;
        return s.substr(b, 1);
    }
    void expect(const char *t) {
        const std::string w = next();
        if (w != t)
            throw std::runtime_error(std::string("SDF header: expected '") + t + "', found '" + w + "'");
    }
};

void parse(SdfFile &F, const std::string &hdr) {
    Lexer L(hdr);
    for (;;) {
        std::string w = L.next();
        if (w.empty())
            break;
        if (w == "parameter") {
            const std::string name = L.next();
            L.expect("=");
            const std::string v = L.next();
            L.expect(";");
            if (name == "byteorder" && strtoull(v.c_str(), nullptr, 0) != 0x78563412ull)
                throw std::runtime_error("SDF: byteorder " + v + " (only little-endian files are read)");
            continue;
        }
        if (w == "struct") {
            L.expect("{");
            int64_t off = 0;
            for (;;) {
                std::string t = L.next();
                if (t == "}")
                    break;
                int ty;
                if (!type_of(t, &ty))
                    throw std::runtime_error("SDF struct: unknown member type '" + t + "'");
                for (;;) {
                    const std::string name = L.next();
                    std::string sep = L.next();
                    int64_t count = 1;
                    if (sep == "[") {
                        count = strtoll(L.next().c_str(), nullptr, 10);
                        L.expect("]");
                        sep = L.next();
                    }
                    for (int64_t c = 0; c < count; c++) {
                        F.fields.push_back({count > 1 ? name + "[" + std::to_string(c) + "]" : name, ty, off});
                        off += type_size(ty);
                    }
                    if (sep == ";")
                        break;
                    if (sep != ",")
                        throw std::runtime_error("SDF struct: expected ',' or ';' after " + name);
                }
            }
            F.recsize = off;
            L.expect("[");
            std::string n = L.next();
            if (n == "]") {
                F.nrec = -1;  // from the file size
            } else {
                F.nrec = strtoll(n.c_str(), nullptr, 10);
                L.expect("]");
            }
            L.expect(";");
            continue;
        }
        int ty;
        if (!type_of(w, &ty))
            throw std::runtime_error("SDF header: unknown declaration '" + w + "'");
        Value v;
        v.name = L.next();
        std::string t = L.next();
        if (t == "[") {  // char name[] = "...";
            L.expect("]");
            L.expect("=");
            const std::string s = L.next();
            if (s.size() < 2 || s[0] != '"')
                throw std::runtime_error("SDF header: string value expected for " + v.name);
            v.type = PGRAV_SDF_STRING;
            v.s = s.substr(1, s.size() - 2);
        } else {
            if (t != "=")
                throw std::runtime_error("SDF header: expected '=' after " + v.name);
            const std::string num = L.next();
            v.type = ty;
            if (ty == PGRAV_SDF_FLOAT || ty == PGRAV_SDF_DOUBLE)
                v.d = strtod(num.c_str(), nullptr);
            else
                v.i = strtoll(num.c_str(), nullptr, 0);
        }
        L.expect(";");
        F.values.push_back(v);
    }
}

}  // namespace

}  // namespace pgrav

using namespace pgrav;

extern "C" {

int pgrav_sdf_write(const char *path, const pgrav_sdf_value *vals, int nvals, const char *desc,
                    const void *records, int64_t nrec, int64_t recsize) {
    clear_error();
    if (!path || (nvals && !vals) || !desc || nrec < 0 || recsize <= 0 || (nrec && !records)) {
        fail("pgrav_sdf_write: bad arguments");
        return -1;
    }
    std::string h = "# SDF\nparameter byteorder = 0x78563412;\n";
    char line[4096];
    for (int k = 0; k < nvals; k++) {
        const pgrav_sdf_value &v = vals[k];
        if (!v.name) {
            fail("pgrav_sdf_write: value %d has no name", k);
            return -1;
        }
        switch (v.type) {
        case PGRAV_SDF_INT: snprintf(line, sizeof line, "int %s = %d;\n", v.name, (int)v.i); break;
        case PGRAV_SDF_INT64: snprintf(line, sizeof line, "int64_t %s = %lld;\n", v.name, (long long)v.i); break;
        case PGRAV_SDF_FLOAT: snprintf(line, sizeof line, "float %s = %g;\n", v.name, v.d); break;
        case PGRAV_SDF_DOUBLE: snprintf(line, sizeof line, "double %s = %.16g;\n", v.name, v.d); break;
        case PGRAV_SDF_STRING:
            snprintf(line, sizeof line, "char %s[] = \"%s\";\n", v.name, v.s ? v.s : "");
            break;
        default: fail("pgrav_sdf_write: value %s has an unknown type %d", v.name, v.type); return -1;
        }
        h += line;
    }
    h += desc;
    snprintf(line, sizeof line, "[%lld];\n", (long long)nrec);
    h += line;
    h += "#\f\n# SDF-EOH ";
    // pad so that the data begins on a DATAALIGN boundary (sph+nln/wind.c:128-138)
    int pad = (int)((h.size() + 1) % DATAALIGN);
    if (pad)
        pad = DATAALIGN - pad;
    h.append(pad, ' ');
    h += '\n';
    FILE *f = fopen(path, "wb");
    if (!f) {
        fail("pgrav_sdf_write: cannot open %s", path);
        return -1;
    }
    bool ok = fwrite(h.data(), 1, h.size(), f) == h.size();
    if (ok && nrec)
        ok = fwrite(records, (size_t)recsize, (size_t)nrec, f) == (size_t)nrec;
    ok = (fclose(f) == 0) && ok;
    if (!ok) {
        fail("pgrav_sdf_write: write to %s failed", path);
        return -1;
    }
    return 0;
}

void *pgrav_sdf_open(const char *path) {
    clear_error();
    if (!path) {
        fail("pgrav_sdf_open: null path");
        return nullptr;
    }
    FILE *f = fopen(path, "rb");
    if (!f) {
        fail("pgrav_sdf_open: cannot open %s", path);
        return nullptr;
    }
    std::string hdr;
    char buf[4096];
    bool eoh = false;
    int64_t pos = 0;
    while (fgets(buf, sizeof buf, f)) {
        const size_t len = strlen(buf);
        pos += (int64_t)len;
        if (strncmp(buf, "# SDF-EOH", 9) == 0) {
            eoh = true;
            break;
        }
        hdr += buf;
        if (hdr.size() > (64u << 20))
            break;
    }
    fseek(f, 0, SEEK_END);
    const int64_t size = ftell(f);
    fclose(f);
    if (!eoh || hdr.compare(0, 5, "# SDF") != 0) {
        fail("pgrav_sdf_open: %s is not an SDF file (no '# SDF' ... '# SDF-EOH' header)", path);
        return nullptr;
    }
    std::unique_ptr<SdfFile> F(new SdfFile());
    F->path = path;
    try {
        parse(*F, hdr);
    } catch (const std::exception &e) {
        fail("pgrav_sdf_open: %s: %s", path, e.what());
        return nullptr;
    }
    F->data_offset = pos;
    if (F->recsize > 0) {
        const int64_t avail = (size - pos) / F->recsize;
        if (F->nrec < 0)
            F->nrec = avail;
        if (F->nrec > avail) {
            fail("pgrav_sdf_open: %s declares %lld records, holds %lld", path, (long long)F->nrec,
                 (long long)avail);
            return nullptr;
        }
    } else {
        F->nrec = 0;
    }
    return F.release();
}

void pgrav_sdf_close(void *h) { delete (SdfFile *)h; }

int pgrav_sdf_nvalues(void *h) { return h ? (int)((SdfFile *)h)->values.size() : -1; }

int pgrav_sdf_value_at(void *h, int k, pgrav_sdf_value *out) {
    SdfFile *F = (SdfFile *)h;
    if (!F || !out || k < 0 || k >= (int)F->values.size()) {
        fail("pgrav_sdf_value_at: bad arguments");
        return -1;
    }
    const Value &v = F->values[k];
    out->name = v.name.c_str();
    out->type = v.type;
    out->i = v.i;
    out->d = v.d;
    out->s = v.s.c_str();
    return 0;
}

int64_t pgrav_sdf_nrecords(void *h) { return h ? ((SdfFile *)h)->nrec : -1; }

int64_t pgrav_sdf_recsize(void *h) { return h ? ((SdfFile *)h)->recsize : -1; }

int pgrav_sdf_nfields(void *h) { return h ? (int)((SdfFile *)h)->fields.size() : -1; }

int pgrav_sdf_field_at(void *h, int k, const char **name, int *type, int64_t *offset) {
    SdfFile *F = (SdfFile *)h;
    if (!F || k < 0 || k >= (int)F->fields.size()) {
        fail("pgrav_sdf_field_at: bad arguments");
        return -1;
    }
    if (name)
        *name = F->fields[k].name.c_str();
    if (type)
        *type = F->fields[k].type;
    if (offset)
        *offset = F->fields[k].offset;
    return 0;
}

int pgrav_sdf_read(void *h, int64_t first, int64_t count, void *out) {
    SdfFile *F = (SdfFile *)h;
    clear_error();
    if (!F || first < 0 || count < 0 || first + count > F->nrec || (count && !out)) {
        fail("pgrav_sdf_read: bad record range");
        return -1;
    }
    FILE *f = fopen(F->path.c_str(), "rb");
    if (!f) {
        fail("pgrav_sdf_read: cannot open %s", F->path.c_str());
        return -1;
    }
    bool ok = fseek(f, (long)(F->data_offset + first * F->recsize), SEEK_SET) == 0 &&
              fread(out, (size_t)F->recsize, (size_t)count, f) == (size_t)count;
    fclose(f);
    if (!ok) {
        fail("pgrav_sdf_read: short read from %s", F->path.c_str());
        return -1;
    }
    return 0;
}

}  // extern "C"
