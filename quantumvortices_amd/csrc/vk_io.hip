// Checkpoint formats of the reference, host side.
//
// wf files (src/io_fields.f90:41-50, restart_gpe :71-224, and
// restart_from_argle_data src/initial_config.f90:630-771): one Fortran
// unformatted sequential file per slab, 'wf<iopoint>p<rank+1>.dat', holding
// a header record (INTEGER iopoint, REAL(8) time, REAL(8) dt) and then one
// record of Nx COMPLEX(8) values per local row y.  Each record is framed by
// its byte length as a 4-byte little-endian integer before and after (the
// gfortran/flang sequential layout).  Readers regroup nsplit_orignal slabs
// of Ny/nsplit_orignal rows into one field (the nsplit_orignal >
// nsplit_current branch with nsplit_current = 1).
//
// Particle files (src/gpe_particle.f90:92-184): list-directed text, one line
// per particle 'iopoint ptime a b' in gppos / gpvel / gpf0ext files;
// checkpoints append, restarts read the first np lines.
#include <hip/hip_runtime.h>

#include <cerrno>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "vikshobha.h"

int vk_set_error(int code, const std::string& msg);   // vk_gpe.hip (vk_last_error)

namespace {

std::string join(const char* dir, const std::string& name)
{
    std::string d = dir ? dir : ".";
    if (!d.empty() && d.back() != '/') d += '/';
    return d + name;
}

std::string wf_name(int iopoint, int rank1) { return "wf" + std::to_string(iopoint) + "p" + std::to_string(rank1) + ".dat"; }

struct File {
    FILE* f = nullptr;
    ~File()
    {
        if (f) fclose(f);
    }
};

bool put_record(FILE* f, const void* p, size_t n)
{
    const int32_t m = (int32_t)n;
    return fwrite(&m, 4, 1, f) == 1 && (n == 0 || fwrite(p, 1, n, f) == n) && fwrite(&m, 4, 1, f) == 1;
}

// reads one record of exactly n bytes into p
int get_record(FILE* f, void* p, size_t n, const std::string& what)
{
    int32_t m0 = 0, m1 = 0;
    if (fread(&m0, 4, 1, f) != 1) return vk_set_error(VK_ERR_ARG, what + ": unexpected end of file");
    if (m0 < 0 || (size_t)m0 != n)
        return vk_set_error(VK_ERR_ARG, what + ": record of " + std::to_string(m0) + " bytes, expected " +
                                            std::to_string(n));
    if (fread(p, 1, n, f) != n || fread(&m1, 4, 1, f) != 1 || m1 != m0)
        return vk_set_error(VK_ERR_ARG, what + ": truncated or inconsistent record");
    return 0;
}

#pragma pack(push, 1)
struct WfHeader {
    int32_t iopoint;
    double time;
    double dt;
};
#pragma pack(pop)
static_assert(sizeof(WfHeader) == 20, "INTEGER + 2 REAL(8) without padding");

// one list-directed line 'iopoint ptime a b'; Fortran may write D exponents
bool parse_line(const char* line, double* a, double* b)
{
    std::string s(line);
    for (char& c : s)
        if (c == 'D' || c == 'd') c = 'E';
    for (char& c : s)
        if (c == ',') c = ' ';
    char* e = nullptr;
    const char* p = s.c_str();
    strtol(p, &e, 10);
    if (e == p) return false;
    p = e;
    strtod(p, &e);
    if (e == p) return false;
    p = e;
    *a = strtod(p, &e);
    if (e == p) return false;
    p = e;
    *b = strtod(p, &e);
    return e != p;
}

int read_pairs(const std::string& path, int np, double* out)
{
    File fh;
    fh.f = fopen(path.c_str(), "r");
    if (!fh.f) return vk_set_error(VK_ERR_ARG, "cannot open " + path + ": " + strerror(errno));
    char line[1024];
    for (int i = 0; i < np; ++i) {
        if (!fgets(line, sizeof line, fh.f)) return vk_set_error(VK_ERR_ARG, path + ": fewer than np lines");
        if (!parse_line(line, out + 2 * i, out + 2 * i + 1))
            return vk_set_error(VK_ERR_ARG, path + ": cannot parse line " + std::to_string(i + 1));
    }
    return 0;
}

int append_pairs(const std::string& path, int iopoint, double ptime, int np, const double* v)
{
    File fh;
    fh.f = fopen(path.c_str(), "a");
    if (!fh.f) return vk_set_error(VK_ERR_ARG, "cannot open " + path + ": " + strerror(errno));
    for (int i = 0; i < np; ++i)
        if (fprintf(fh.f, "%12d %25.17E %25.17E %25.17E\n", iopoint, ptime, v[2 * i], v[2 * i + 1]) < 0)
            return vk_set_error(VK_ERR_ARG, path + ": write failed");
    return 0;
}

}  // namespace

extern "C" {

int vk_wf_write(const char* dir, int iopoint, double time, double dt, int N, int nsplit, const double* psi)
{
    if (!psi || N < 1 || nsplit < 1 || N % nsplit) return vk_set_error(VK_ERR_ARG, "wf_write: bad arguments");
    const int rows = N / nsplit;
    const WfHeader hd{iopoint, time, dt};
    for (int r = 0; r < nsplit; ++r) {
        const std::string path = join(dir, wf_name(iopoint, r + 1));
        File fh;
        fh.f = fopen(path.c_str(), "wb");
        if (!fh.f) return vk_set_error(VK_ERR_ARG, "cannot create " + path + ": " + strerror(errno));
        if (!put_record(fh.f, &hd, sizeof hd)) return vk_set_error(VK_ERR_ARG, path + ": write failed");
        for (int y = r * rows; y < (r + 1) * rows; ++y)
            if (!put_record(fh.f, psi + 2 * (size_t)y * N, 16 * (size_t)N))
                return vk_set_error(VK_ERR_ARG, path + ": write failed");
    }
    return 0;
}

int vk_wf_read(const char* dir, int nrestart, int N, int nsplit_orig, int* iopoint, double* time, double* dt,
               double* psi)
{
    if (!psi || N < 1 || nsplit_orig < 1 || N % nsplit_orig) return vk_set_error(VK_ERR_ARG, "wf_read: bad arguments");
    const int rows = N / nsplit_orig;
    for (int r = 0; r < nsplit_orig; ++r) {
        const std::string path = join(dir, wf_name(nrestart, r + 1));
        File fh;
        fh.f = fopen(path.c_str(), "rb");
        if (!fh.f) return vk_set_error(VK_ERR_ARG, "cannot open " + path + ": " + strerror(errno));
        WfHeader hd{};
        int rc = get_record(fh.f, &hd, sizeof hd, path + " header");
        if (rc) return rc;
        if (r == 0) {   // the last file read sets them in the reference; all files carry the same values
            if (iopoint) *iopoint = hd.iopoint;
            if (time) *time = hd.time;
            if (dt) *dt = hd.dt;
        }
        for (int y = r * rows; y < (r + 1) * rows; ++y)
            if ((rc = get_record(fh.f, psi + 2 * (size_t)y * N, 16 * (size_t)N, path + " row " + std::to_string(y + 1))))
                return rc;
    }
    return 0;
}

int vk_part_checkpoint(const char* dir, int nrst, int iopoint, double ptime, int np, const double* pos,
                       const double* vel, const double* fext)
{
    if (np < 0 || (np && (!pos || !vel || !fext))) return vk_set_error(VK_ERR_ARG, "part_checkpoint: bad arguments");
    const std::string sfx = ".dat_nrst" + std::to_string(nrst);
    int rc;
    if ((rc = append_pairs(join(dir, "gppos_ckp" + sfx), iopoint, ptime, np, pos))) return rc;
    if ((rc = append_pairs(join(dir, "gpvel_ckp" + sfx), iopoint, ptime, np, vel))) return rc;
    return append_pairs(join(dir, "gpf0ext_ckp" + sfx), iopoint, ptime, np, fext);
}

int vk_part_restart(const char* dir, int nrst, int np, double* pos, double* vel, double* fext)
{
    if (np < 0 || (np && (!pos || !vel || !fext))) return vk_set_error(VK_ERR_ARG, "part_restart: bad arguments");
    const std::string sfx = ".dat_nrst" + std::to_string(nrst);
    int rc;
    if ((rc = read_pairs(join(dir, "gppos" + sfx), np, pos))) return rc;
    if ((rc = read_pairs(join(dir, "gpvel" + sfx), np, vel))) return rc;
    return read_pairs(join(dir, "gpf0ext" + sfx), np, fext);
}

}  // extern "C"
