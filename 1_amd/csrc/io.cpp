// OVF 1.0 / 2.0, .tensor and ODT files (include/mumax_io.h), host C++.
//
// Restates mumax/omf/input.go:21-235 (header keys, "# Desc:" entries,
// "Begin: Data Text|Binary 4", control number 1234567.0, nodes with the user
// x index fastest and the user (x, y, z) components per node, i.e. internal
// components Z, Y, X), mumax/omf/output.go:37-300 (OVF 1.0: big-endian
// Binary 4; OVF 2.0: little-endian Binary 4, cell-centred bases, valuedim 3),
// mumax/tensor/{input,output}.go (magic 0x0A317423 = "#t1\n", int32 rank and
// sizes, float32 data) and mumax/omf/odt.go:26-132 (ODT 1.0 header, tab
// separated rows, "# Table End").  Reading also accepts Binary 8 (OOMMF's
// control number 123456789012345.0) and either byte order.
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mumax_io.h"

namespace {

thread_local std::string g_err;

int fail(const std::string& msg) {
    g_err = msg;
    return -1;
}

std::string lower(std::string s) {
    for (auto& c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}
std::string trim(const std::string& s, const char* set = " \t\r\n#") {
    const size_t a = s.find_first_not_of(set);
    if (a == std::string::npos) return "";
    const size_t b = s.find_last_not_of(set);
    return s.substr(a, b - a + 1);
}
void copy_str(char* dst, const std::string& s, size_t cap) {
    if (!dst || !cap) return;
    std::strncpy(dst, s.c_str(), cap - 1);
    dst[cap - 1] = 0;
}

struct Header {
    mmOvfInfo info;
    std::vector<std::pair<std::string, std::string>> desc;
    std::string format;  // "text" / "binary"
    int bytes = 0;
    long data_offset = 0;
};

// "# key: value" up to "# Begin: Data <format> [<n>]"
int read_header(FILE* f, Header& h) {
    std::memset(&h.info, 0, sizeof(h.info));
    h.info.multiplier = 1.0;
    h.info.version = 1;
    char line[4096];
    while (std::fgets(line, sizeof line, f)) {
        std::string s = line;
        const std::string t = trim(s);
        if (t.empty()) continue;  // "#" separator lines (OVF 2.0)
        const size_t colon = t.find(':');
        if (colon == std::string::npos) continue;
        const std::string key = lower(trim(t.substr(0, colon)));
        const std::string value = trim(t.substr(colon + 1), " \t\r\n");
        if (key == "begin") {
            std::string v = lower(value);
            if (v.rfind("data", 0) == 0) {
                const std::string rest = trim(v.substr(4), " \t");
                if (rest.rfind("text", 0) == 0) {
                    h.format = "text";
                    h.info.binary = 0;
                } else if (rest.rfind("binary", 0) == 0) {
                    h.format = "binary";
                    h.bytes = std::atoi(rest.substr(6).c_str());
                    if (h.bytes != 4 && h.bytes != 8) return fail("unsupported data format: " + value);
                    h.info.binary = h.bytes;
                } else {
                    return fail("unknown data format: " + value);
                }
                h.data_offset = std::ftell(f);
                return 0;
            }
        } else if (key == "oommf") {
            h.info.version = lower(value).find("2.0") != std::string::npos ? 2 : 1;
        } else if (key == "xnodes") {
            h.info.size[2] = std::atoi(value.c_str());
        } else if (key == "ynodes") {
            h.info.size[1] = std::atoi(value.c_str());
        } else if (key == "znodes") {
            h.info.size[0] = std::atoi(value.c_str());
        } else if (key == "xstepsize") {
            h.info.step[2] = std::atof(value.c_str());
        } else if (key == "ystepsize") {
            h.info.step[1] = std::atof(value.c_str());
        } else if (key == "zstepsize") {
            h.info.step[0] = std::atof(value.c_str());
        } else if (key == "valuemultiplier") {
            h.info.multiplier = std::atof(value.c_str());
        } else if (key == "valueunit" || key == "valueunits") {
            std::string u = value;
            const size_t sp = u.find_first_of(" \t");
            if (sp != std::string::npos) u = u.substr(0, sp);
            copy_str(h.info.valueunit, u, sizeof h.info.valueunit);
        } else if (key == "meshunit") {
            copy_str(h.info.meshunit, value, sizeof h.info.meshunit);
        } else if (key == "title") {
            copy_str(h.info.title, value, sizeof h.info.title);
        } else if (key == "desc") {
            const size_t c2 = value.find(':');
            std::string dk = trim(c2 == std::string::npos ? value : value.substr(0, c2), " \t#");
            std::string dv = c2 == std::string::npos ? "" : trim(value.substr(c2 + 1), " \t#");
            const std::string lk = lower(dk);
            if (lk == "total simulation time") h.info.total_time = std::atof(dv.c_str());
            if (lk == "stage simulation time") h.info.stage_time = std::atof(dv.c_str());
            h.desc.emplace_back(dk, dv);
        }
    }
    return fail("no \"Begin: Data\" line");
}

int open_header(const char* path, FILE** fp, Header& h) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(std::string("cannot open ") + path);
    if (read_header(f, h)) {
        std::fclose(f);
        return -1;
    }
    h.info.ndesc = (int)h.desc.size();
    if (h.info.size[0] <= 0 || h.info.size[1] <= 0 || h.info.size[2] <= 0) {
        std::fclose(f);
        return fail("missing or bad xnodes / ynodes / znodes");
    }
    *fp = f;
    return 0;
}

uint32_t bswap32(uint32_t v) { return (v >> 24) | ((v >> 8) & 0xff00) | ((v << 8) & 0xff0000) | (v << 24); }
uint64_t bswap64(uint64_t v) { return ((uint64_t)bswap32((uint32_t)v) << 32) | bswap32((uint32_t)(v >> 32)); }

}  // namespace

extern "C" {

const char* mm_io_error(void) { return g_err.c_str(); }

int mm_ovf_info(const char* path, mmOvfInfo* info) {
    FILE* f = nullptr;
    Header h;
    if (open_header(path, &f, h)) return -1;
    std::fclose(f);
    if (info) *info = h.info;
    return 0;
}

int mm_ovf_desc(const char* path, int i, char* key, char* value, int cap) {
    FILE* f = nullptr;
    Header h;
    if (open_header(path, &f, h)) return -1;
    std::fclose(f);
    if (i < 0 || i >= (int)h.desc.size()) return fail("no Desc entry " + std::to_string(i));
    copy_str(key, h.desc[i].first, cap);
    copy_str(value, h.desc[i].second, cap);
    return 0;
}

int mm_ovf_read(const char* path, float* m, mmOvfInfo* info) {
    FILE* f = nullptr;
    Header h;
    if (open_header(path, &f, h)) return -1;
    const int X = h.info.size[0], Y = h.info.size[1], Z = h.info.size[2];
    const long N = (long)X * Y * Z;
    auto put = [&](long node, int comp_user, double v) {
        const int k = (int)(node % Z), j = (int)((node / Z) % Y), i = (int)(node / ((long)Z * Y));
        const int c = 2 - comp_user;  // user x, y, z -> internal Z, Y, X
        m[((long)c * X + i) * Y * Z + (long)j * Z + k] = (float)v;
    };
    int rc = 0;
    if (h.format == "text") {
        for (long n = 0; n < N && !rc; n++)
            for (int cu = 0; cu < 3; cu++) {
                double v;
                if (std::fscanf(f, "%lf", &v) != 1) {
                    rc = fail("text data ends early at node " + std::to_string(n));
                    break;
                }
                put(n, cu, v);
            }
    } else if (h.bytes == 4) {
        uint32_t w;
        if (std::fread(&w, 4, 1, f) != 1) rc = fail("no control number");
        float le, be;
        std::memcpy(&le, &w, 4);
        const uint32_t ws = bswap32(w);
        std::memcpy(&be, &ws, 4);
        const bool big = be == 1234567.0f;
        if (!rc && !big && le != 1234567.0f) rc = fail("bad control number");
        std::vector<uint32_t> buf(3 * N);
        if (!rc && std::fread(buf.data(), 4, buf.size(), f) != buf.size()) rc = fail("binary data ends early");
        for (long n = 0; n < N && !rc; n++)
            for (int cu = 0; cu < 3; cu++) {
                uint32_t u = buf[3 * n + cu];
                if (big) u = bswap32(u);
                float v;
                std::memcpy(&v, &u, 4);
                put(n, cu, v);
            }
    } else {
        uint64_t w;
        if (std::fread(&w, 8, 1, f) != 1) rc = fail("no control number");
        double le, be;
        std::memcpy(&le, &w, 8);
        const uint64_t ws = bswap64(w);
        std::memcpy(&be, &ws, 8);
        const bool big = be == 123456789012345.0;
        if (!rc && !big && le != 123456789012345.0) rc = fail("bad control number");
        std::vector<uint64_t> buf(3 * N);
        if (!rc && std::fread(buf.data(), 8, buf.size(), f) != buf.size()) rc = fail("binary data ends early");
        for (long n = 0; n < N && !rc; n++)
            for (int cu = 0; cu < 3; cu++) {
                uint64_t u = buf[3 * n + cu];
                if (big) u = bswap64(u);
                double v;
                std::memcpy(&v, &u, 8);
                put(n, cu, v);
            }
    }
    std::fclose(f);
    if (!rc && info) *info = h.info;
    return rc;
}

int mm_ovf_write(const char* path, const float* m, const mmOvfInfo* info, const char* const* desc_keys,
                 const char* const* desc_values, int ndesc) {
    const int X = info->size[0], Y = info->size[1], Z = info->size[2];
    if (X <= 0 || Y <= 0 || Z <= 0) return fail("bad size");
    if (info->version != 1 && info->version != 2) return fail("version must be 1 or 2");
    if (info->binary != 0 && info->binary != 4) return fail("write: text or Binary 4");
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(std::string("cannot create ") + path);
    const double dx = info->step[2], dy = info->step[1], dz = info->step[0];
    const char* unit = info->valueunit[0] ? info->valueunit : "A/m";
    const char* mesh = info->meshunit[0] ? info->meshunit : "m";
    const char* title = info->title[0] ? info->title : "mumax data";
    auto desc = [&]() {
        for (int i = 0; i < ndesc; i++) std::fprintf(f, "# Desc: %s :  %s\n", desc_keys[i], desc_values[i]);
    };
    if (info->version == 1) {
        std::fprintf(f, "# OOMMF: rectangular mesh v1.0\n# Segment count: 1\n# Begin: Segment\n# Begin: Header\n");
        desc();
        std::fprintf(f, "# Title: %s\n# meshtype: rectangular\n# meshunit: %s\n", title, mesh);
        std::fprintf(f, "# xbase: 0\n# ybase: 0\n# zbase: 0\n");
        std::fprintf(f, "# xstepsize: %.9g\n# ystepsize: %.9g\n# zstepsize: %.9g\n", dx, dy, dz);
        std::fprintf(f, "# xmin: 0\n# ymin: 0\n# zmin: 0\n");
        std::fprintf(f, "# xmax: %.9g\n# ymax: %.9g\n# zmax: %.9g\n", dx * Z, dy * Y, dz * X);
        std::fprintf(f, "# xnodes: %d\n# ynodes: %d\n# znodes: %d\n", Z, Y, X);
        std::fprintf(f, "# ValueRangeMinMag: 1e-08\n# ValueRangeMaxMag: 1\n");
        std::fprintf(f, "# valueunit: %s\n# valuemultiplier: %.9g\n# End: Header\n", unit,
                     info->multiplier != 0 ? info->multiplier : 1.0);
    } else {
        std::fprintf(f, "# OOMMF: OVF 2.0\n#\n# Segment count: 1\n#\n# Begin: Segment\n# Begin: Header\n#\n");
        std::fprintf(f, "# title: %s\n# meshtype: rectangular\n# meshunit: %s\n", title, mesh);
        std::fprintf(f, "# xmin: 0\n# ymin: 0\n# zmin: 0\n");
        std::fprintf(f, "# xmax: %.9g\n# ymax: %.9g\n# zmax: %.9g\n", dx * Z, dy * Y, dz * X);
        std::fprintf(f, "# valuedim: 3\n# valuelabels: Val_x Val_y Val_z\n# valueunits: %s %s %s\n", unit, unit, unit);
        desc();
        std::fprintf(f, "# desc: Stage simulation time: %.9g s\n# desc: Total simulation time: %.9g s\n",
                     info->stage_time, info->total_time);
        std::fprintf(f, "# xbase: %.9g\n# ybase: %.9g\n# zbase: %.9g\n", dx / 2, dy / 2, dz / 2);
        std::fprintf(f, "# xnodes: %d\n# ynodes: %d\n# znodes: %d\n", Z, Y, X);
        std::fprintf(f, "# xstepsize: %.9g\n# ystepsize: %.9g\n# zstepsize: %.9g\n#\n# End: Header\n#\n", dx, dy, dz);
    }
    auto value = [&](int i, int j, int k, int cu) { return m[((long)(2 - cu) * X + i) * Y * Z + (long)j * Z + k]; };
    if (info->binary == 0) {
        std::fprintf(f, "# Begin: Data Text\n");
        for (int i = 0; i < X; i++) {
            for (int j = 0; j < Y; j++) {
                for (int k = 0; k < Z; k++)
                    std::fprintf(f, "%.9g %.9g %.9g \t", value(i, j, k, 0), value(i, j, k, 1), value(i, j, k, 2));
                std::fprintf(f, "\n");
            }
            std::fprintf(f, "\n");
        }
        std::fprintf(f, "# End: Data Text\n");
    } else {
        std::fprintf(f, "# Begin: Data Binary 4\n");
        const bool big = info->version == 1;
        auto w32 = [&](float v) {
            uint32_t u;
            std::memcpy(&u, &v, 4);
            if (big) u = bswap32(u);
            std::fwrite(&u, 4, 1, f);
        };
        w32(1234567.0f);
        for (int i = 0; i < X; i++)
            for (int j = 0; j < Y; j++)
                for (int k = 0; k < Z; k++)
                    for (int cu = 0; cu < 3; cu++) w32(value(i, j, k, cu));
        std::fprintf(f, "# End: Data Binary 4\n");
    }
    std::fprintf(f, "# End: Segment\n");
    const bool ok = std::ferror(f) == 0;
    std::fclose(f);
    return ok ? 0 : fail(std::string("write error on ") + path);
}

int mm_tensor_read(const char* path, int* rank, int* size, float* data) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(std::string("cannot open ") + path);
    int32_t hdr[2];
    int rc = 0;
    if (std::fread(hdr, 4, 2, f) != 2) rc = fail("short tensor header");
    if (!rc && hdr[0] != 0x0A317423) rc = fail("bad tensor magic");
    if (!rc && (hdr[1] < 0 || hdr[1] > 16)) rc = fail("bad tensor rank");
    std::vector<int32_t> sz(rc ? 0 : hdr[1]);
    if (!rc && std::fread(sz.data(), 4, sz.size(), f) != sz.size()) rc = fail("short tensor header");
    long n = 1;
    if (!rc) {
        *rank = hdr[1];
        for (int i = 0; i < hdr[1]; i++) {
            size[i] = sz[i];
            n *= sz[i];
        }
        if (data && std::fread(data, 4, n, f) != (size_t)n) rc = fail("tensor data ends early");
    }
    std::fclose(f);
    return rc;
}

int mm_tensor_write(const char* path, int rank, const int* size, const float* data) {
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(std::string("cannot create ") + path);
    int32_t hdr[2] = {0x0A317423, rank};
    std::fwrite(hdr, 4, 2, f);
    long n = 1;
    for (int i = 0; i < rank; i++) {
        const int32_t s = size[i];
        std::fwrite(&s, 4, 1, f);
        n *= size[i];
    }
    std::fwrite(data, 4, n, f);
    const bool ok = std::ferror(f) == 0;
    std::fclose(f);
    return ok ? 0 : fail(std::string("write error on ") + path);
}

}  // extern "C"

struct mmOdt {
    FILE* f = nullptr;
    int ncols = 0;
};

extern "C" {

mmOdt* mm_odt_open(const char* path, const char* title, int ncols, const char* const* names,
                   const char* const* units) {
    return mm_odt_open_desc(path, title, ncols, names, units, 0, nullptr, nullptr);
}

mmOdt* mm_odt_open_desc(const char* path, const char* title, int ncols, const char* const* names,
                        const char* const* units, int ndesc, const char* const* desc_keys,
                        const char* const* desc_values) {
    FILE* f = std::fopen(path, "w");
    if (!f) {
        fail(std::string("cannot create ") + path);
        return nullptr;
    }
    mmOdt* t = new mmOdt;
    t->f = f;
    t->ncols = ncols;
    std::fprintf(f, "# ODT 1.0\n");
    if (ndesc > 0) {
        std::fprintf(f, "# Begin: Header\n");
        for (int i = 0; i < ndesc; i++) std::fprintf(f, "# Desc: %s :  %s\n", desc_keys[i], desc_values[i]);
        std::fprintf(f, "# End: Header\n");
    }
    std::fprintf(f, "# Table Start\n# Title:  %s\n# Units:", title ? title : "");
    for (int i = 0; i < ncols; i++) std::fprintf(f, "{%s} \t", units[i]);
    std::fprintf(f, "\n# Columns:");
    for (int i = 0; i < ncols; i++) std::fprintf(f, "%s \t", names[i]);
    std::fprintf(f, "\n");
    return t;
}

int mm_odt_row(mmOdt* t, const double* values) {
    for (int i = 0; i < t->ncols; i++) std::fprintf(t->f, "%-20.10g \t", values[i]);
    std::fprintf(t->f, "\n");
    std::fflush(t->f);
    return std::ferror(t->f) ? fail("write error") : 0;
}

int mm_odt_close(mmOdt* t) {
    if (!t) return 0;
    std::fprintf(t->f, "# Table End\n");
    const bool ok = std::ferror(t->f) == 0;
    std::fclose(t->f);
    delete t;
    return ok ? 0 : fail("write error");
}

}  // extern "C"
