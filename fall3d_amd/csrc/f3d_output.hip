// Output side of a step (printres_nc.f90, wrirst.f90, rearst.f90), on a
// context's device state:
//
// * the concentration outputs C_GRND, CON (terrain following or a.s.l.)
//   and the flight-level cuts C_FLxxx, one thread per column;
// * the restart file of wrirst.f90 -- the state c with its ghost frame and
//   the deposit load cload(nx,ny,0:nc), both as 32-bit floats, with the
//   run's clock and extent as global attributes -- written and read in the
//   netCDF classic format by the code below (no netCDF library), so that a
//   netCDF reader, rearst.f90 among them, reads what f3d_write_restart
//   writes and the other way round.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "f3d_ctx.h"

bool check_ctx(const f3d_ctx *c);

using namespace f3d;

namespace {

// ---------------- concentration outputs ----------------

// setcut.f90:61-107, ICAO standard atmosphere (three layers): p (Pa) and
// T (K) at height z (m a.s.l.)
void standard_atm(double z, double *p, double *T)
{
    const double goo = 9.80665, Roo = 8.31432, Moo = 0.0289644;
    const double zoo = 0.0, Too = 288.15, poo = 101325.0, Loo = -0.0065;
    const double zo1 = 11000.0, To1 = 216.65, po1 = 22632.0;
    const double zo2 = 20000.0, To2 = 216.65, po2 = 5474.0, Lo2 = 0.001;
    if (z < zo1) {
        *T = Too + Loo * (z - zoo);
        *p = poo * std::pow(Too / *T, goo * Moo / (Roo * Loo));
    } else if (z < zo2) {
        *T = To1;
        *p = po1 * std::exp(-goo * Moo * (z - zo1) / (Roo * To1));
    } else {
        *T = To2 + Lo2 * (z - zo2);
        *p = po2 * std::pow(To2 / *T, goo * Moo / (Roo * Lo2));
    }
}

struct ConcArgs {
    Grid g;
    const double *hm1, *zlayer, *topg, *p, *pnom;
    int nfl, asl;
    double *cgrnd, *con, *cfl;
};

// class sum of c at level k of column (i,j), in class order
template <typename T>
__device__ __forceinline__ double level_sum(const Grid &g, const T *c, int i, int j, int k)
{
    double s = 0.0;
    for (int q = 0; q < g.nc; ++q) s = s + (double)c[(long)q * g.G + cidx(g.nx, g.ny, i, j, k)];
    return s;
}

// printres_nc.f90:391-396 (C_GRND), :545-560 with setcut.f90:24-57 (the
// flight-level cuts) and :574-586 with ter2asl.f90:25-50 (CON), for one
// member's slabs c; each value in the reference's operation order
template <typename T>
__global__ void k_conc(ConcArgs a, const T *c)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const Grid &g = a.g;
    if (t >= g.NXY) return;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int i = (int)(t % nx) + 1, j = (int)(t / nx) + 1;
    const double h = a.hm1[t];
    if (a.cgrnd) a.cgrnd[t] = 1e3 * level_sum(g, c, i, j, 1) / h;
    if (a.con) {
        if (!a.asl) {
            for (int k = 1; k <= nz; ++k) a.con[(long)(k - 1) * g.NXY + t] = 1e3 * (level_sum(g, c, i, j, k) / h);
        } else {
            // ter2asl: the a.s.l. height zlayer(k) of the column interpolated
            // between the terrain-following levels that bracket it
            const double zo = fmax(0.0, a.topg[t]);
            for (int k = 1; k <= nz; ++k) {
                const double zasl = a.zlayer[k - 1];
                double v = 0.0;
                if (!(zasl < zo + a.zlayer[0]))
                    for (int kk = 1; kk + 1 <= nz; ++kk)
                        if (zasl >= zo + a.zlayer[kk - 1] && zasl <= zo + a.zlayer[kk]) {
                            const double s = (zo + a.zlayer[kk] - zasl) / (a.zlayer[kk] - a.zlayer[kk - 1]);
                            const double lo = 1e3 * (level_sum(g, c, i, j, kk) / h);
                            const double hi = 1e3 * (level_sum(g, c, i, j, kk + 1) / h);
                            v = s * lo + (1.0 - s) * hi;
                            break;
                        }
                a.con[(long)(k - 1) * g.NXY + t] = v;
            }
        }
    }
    for (int l = 0; l < a.nfl; ++l) {
        // setcut_FL: zero below the first level's pressure and above the
        // last bracket, else linear in pressure between the two levels
        const double pn = a.pnom[l];
        double v = 0.0;
        if (!(pn > a.p[midx(nx, ny, i, j, 1)]))
            for (int k = 1; k + 1 <= nz; ++k) {
                const double pk = a.p[midx(nx, ny, i, j, k)], pk1 = a.p[midx(nx, ny, i, j, k + 1)];
                if (pn <= pk && pn > pk1) {
                    const double s = (pn - pk) / (pk1 - pk);
                    v = (level_sum(g, c, i, j, k) / h) * (1.0 - s) + (level_sum(g, c, i, j, k + 1) / h) * s;
                    break;
                }
            }
        a.cfl[(long)l * g.NXY + t] = 1e3 * v;
    }
}

// ---------------- netCDF classic (64-bit offset) ----------------

enum : uint32_t { NC_DIMENSION = 0x0A, NC_VARIABLE = 0x0B, NC_ATTRIBUTE = 0x0C };
enum : uint32_t { NC_BYTE = 1, NC_CHAR = 2, NC_SHORT = 3, NC_INT = 4, NC_FLOAT = 5, NC_DOUBLE = 6 };

struct NcOut {
    std::vector<unsigned char> b;
    void u32(uint32_t v)
    {
        for (int s = 24; s >= 0; s -= 8) b.push_back((unsigned char)(v >> s));
    }
    void u64(uint64_t v)
    {
        for (int s = 56; s >= 0; s -= 8) b.push_back((unsigned char)(v >> s));
    }
    void pad()
    {
        while (b.size() % 4) b.push_back(0);
    }
    void name(const std::string &s)
    {
        u32((uint32_t)s.size());
        b.insert(b.end(), s.begin(), s.end());
        pad();
    }
    void f64(double d)
    {
        uint64_t v;
        std::memcpy(&v, &d, 8);
        u64(v);
    }
};

struct NcAtt {
    std::string name;
    uint32_t type;
    std::string text;        // NC_CHAR
    std::vector<double> num;  // NC_INT / NC_DOUBLE
};

void put_atts(NcOut &o, const std::vector<NcAtt> &atts)
{
    if (atts.empty()) {
        o.u32(0);
        o.u32(0);
        return;
    }
    o.u32(NC_ATTRIBUTE);
    o.u32((uint32_t)atts.size());
    for (const NcAtt &a : atts) {
        o.name(a.name);
        o.u32(a.type);
        if (a.type == NC_CHAR) {
            o.u32((uint32_t)a.text.size());
            o.b.insert(o.b.end(), a.text.begin(), a.text.end());
            o.pad();
        } else {
            o.u32((uint32_t)a.num.size());
            for (double v : a.num) {
                if (a.type == NC_INT)
                    o.u32((uint32_t)(int32_t)v);
                else
                    o.f64(v);
            }
        }
    }
}

struct NcIn {
    const unsigned char *p;
    size_t n, at = 0;
    bool ok = true;
    uint32_t u32()
    {
        if (at + 4 > n) {
            ok = false;
            return 0;
        }
        uint32_t v = 0;
        for (int k = 0; k < 4; ++k) v = (v << 8) | p[at++];
        return v;
    }
    uint64_t u64()
    {
        const uint64_t hi = u32();
        return (hi << 32) | u32();
    }
    void skip(size_t k)
    {
        if (at + k > n) ok = false;
        at = std::min(n, at + k);
    }
    std::string name()
    {
        const uint32_t len = u32();
        if (!ok || at + len > n) {
            ok = false;
            return std::string();
        }
        std::string s((const char *)p + at, len);
        skip((len + 3) / 4 * 4);
        return s;
    }
};

size_t type_size(uint32_t t)
{
    return (t == NC_BYTE || t == NC_CHAR) ? 1 : t == NC_SHORT ? 2 : (t == NC_INT || t == NC_FLOAT) ? 4
           : t == NC_DOUBLE ? 8 : 0;
}

struct NcVarIn {
    std::string name;
    std::vector<uint32_t> dims;
    uint32_t type = 0;
    uint64_t begin = 0;
};

// Parses the header of a netCDF classic file (CDF-1 or CDF-2, no record
// variables are needed here): dimensions, numeric global attributes and
// the variables' shapes and offsets.  Returns false on a malformed header.
bool parse_header(NcIn &in, std::vector<std::pair<std::string, uint64_t>> &dims,
                  std::vector<std::pair<std::string, double>> &gatt, std::vector<NcVarIn> &vars)
{
    if (in.n < 8 || std::memcmp(in.p, "CDF", 3) != 0) return false;
    const int version = in.p[3];
    if (version != 1 && version != 2) return false;
    in.at = 4;
    (void)in.u32();  // numrecs
    auto atts = [&](std::vector<std::pair<std::string, double>> *out) {
        const uint32_t tag = in.u32(), na = in.u32();
        if (tag != NC_ATTRIBUTE && !(tag == 0 && na == 0)) in.ok = false;
        for (uint32_t a = 0; a < na && in.ok; ++a) {
            const std::string nm = in.name();
            const uint32_t t = in.u32(), ne = in.u32();
            const size_t ts = type_size(t);
            if (!ts) {
                in.ok = false;
                break;
            }
            // the first value of a numeric attribute, converted to double
            // as nf90_get_att converts it into a real(rp) or integer
            if (out && ne >= 1 && t != NC_CHAR && in.at + ts <= in.n) {
                const unsigned char *q = in.p + in.at;
                double v = 0.0;
                if (t == NC_BYTE) {
                    v = (double)(int8_t)q[0];
                } else if (t == NC_SHORT) {
                    v = (double)(int16_t)(((unsigned)q[0] << 8) | q[1]);
                } else if (t == NC_INT || t == NC_FLOAT) {
                    const uint32_t u = ((uint32_t)q[0] << 24) | ((uint32_t)q[1] << 16) |
                                       ((uint32_t)q[2] << 8) | q[3];
                    if (t == NC_INT) {
                        v = (double)(int32_t)u;
                    } else {
                        float x;
                        std::memcpy(&x, &u, 4);
                        v = (double)x;
                    }
                } else {
                    uint64_t u = 0;
                    for (int k = 0; k < 8; ++k) u = (u << 8) | q[k];
                    std::memcpy(&v, &u, 8);
                }
                out->push_back({nm, v});
            }
            in.skip((ne * ts + 3) / 4 * 4);
        }
    };
    uint32_t tag = in.u32(), nd = in.u32();
    if (tag != NC_DIMENSION && !(tag == 0 && nd == 0)) return false;
    for (uint32_t d = 0; d < nd && in.ok; ++d) {
        const std::string nm = in.name();
        dims.push_back({nm, in.u32()});
    }
    atts(&gatt);
    tag = in.u32();
    const uint32_t nv = in.u32();
    if (tag != NC_VARIABLE && !(tag == 0 && nv == 0)) return false;
    for (uint32_t v = 0; v < nv && in.ok; ++v) {
        NcVarIn var;
        var.name = in.name();
        const uint32_t ndim = in.u32();
        for (uint32_t k = 0; k < ndim && in.ok; ++k) var.dims.push_back(in.u32());
        atts(nullptr);
        var.type = in.u32();
        (void)in.u32();  // vsize
        var.begin = version == 2 ? in.u64() : in.u32();
        vars.push_back(var);
    }
    return in.ok;
}

}  // namespace

extern "C" {

int f3d_set_pressure(f3d_ctx *c, const double *p)
{
    if (!check_ctx(c)) return F3D_EINVAL;
    if (!p) {
        set_error("f3d_set_pressure: null field");
        return F3D_EINVAL;
    }
    const size_t b = sizeof(double) * c->g.N;
    if (!c->d_p.p && c->d_p.alloc(b, &c->bytes)) {
        set_error("f3d_set_pressure: out of device memory");
        return F3D_ENOMEM;
    }
    F3D_HIP(hipMemcpyAsync(c->d_p.p, p, b, hipMemcpyHostToDevice, c->stream));
    F3D_HIP(hipStreamSynchronize(c->stream));
    c->p_set = true;
    return F3D_OK;
}

int f3d_output_concentration(f3d_ctx *c, int member, double *cgrnd, double *con, int asl,
                             int nfl, const double *fl, double *cfl)
{
    if (!check_ctx(c)) return F3D_EINVAL;
    const Grid &g = c->g;
    if (member < 0 || member >= g.nm) {
        set_error("f3d_output_concentration: member %d out of range (ensemble of %d)", member, g.nm);
        return F3D_EINVAL;
    }
    if (nfl < 0 || (nfl > 0 && (!fl || !cfl))) {
        set_error("f3d_output_concentration: %d flight levels need fl and cfl", nfl);
        return F3D_EINVAL;
    }
    if (nfl > 0 && !c->p_set) {
        set_error("f3d_output_concentration: the flight-level cuts need f3d_set_pressure");
        return F3D_ESTATE;
    }
    if (con && asl && !c->topg_set) {
        set_error("f3d_output_concentration: a.s.l. concentration needs f3d_set_topography");
        return F3D_ESTATE;
    }
    // nominal pressure of each flight level (fl in hundreds of feet)
    std::vector<double> pnom(std::max(nfl, 1));
    for (int l = 0; l < nfl; ++l) {
        double T;
        standard_atm(fl[l] * 30.48, &pnom[l], &T);
    }
    const size_t n2 = (size_t)g.NXY, n3 = (size_t)g.N, nf = n2 * (size_t)nfl;
    DeviceBuffer out, dpn;
    if (out.alloc(sizeof(double) * (n2 + n3 + nf), nullptr) ||
        dpn.alloc(sizeof(double) * pnom.size(), nullptr)) {
        set_error("f3d_output_concentration: out of device memory");
        return F3D_ENOMEM;
    }
    hipStream_t s = c->stream;
    F3D_HIP(hipMemcpyAsync(dpn.p, pnom.data(), sizeof(double) * pnom.size(), hipMemcpyHostToDevice, s));
    double *dg = out.as<double>(), *dc = dg + n2, *df = dc + n3;
    ConcArgs a;
    a.g = g;
    a.hm1 = c->d_hm1.as<double>();
    a.zlayer = c->d_zlayer.as<double>();
    a.topg = c->topg_set ? c->d_topg.as<double>() : nullptr;
    a.p = c->p_set ? c->d_p.as<double>() : nullptr;
    a.pnom = dpn.as<double>();
    a.nfl = nfl;
    a.asl = asl ? 1 : 0;
    a.cgrnd = cgrnd ? dg : nullptr;
    a.con = con ? dc : nullptr;
    a.cfl = nfl > 0 ? df : nullptr;
    const unsigned nb = (unsigned)((g.NXY + 255) / 256);
    const long slab = (long)member * g.nc * g.G;
    if (c->precision == F3D_FP64)
        hipLaunchKernelGGL((k_conc<double>), dim3(nb), dim3(256), 0, s, a, c->d_c64.as<double>() + slab);
    else
        hipLaunchKernelGGL((k_conc<float>), dim3(nb), dim3(256), 0, s, a, c->d_c32[c->cur].as<float>() + slab);
    F3D_CHECK_LAUNCH();
    if (cgrnd) F3D_HIP(hipMemcpyAsync(cgrnd, dg, sizeof(double) * n2, hipMemcpyDeviceToHost, s));
    if (con) F3D_HIP(hipMemcpyAsync(con, dc, sizeof(double) * n3, hipMemcpyDeviceToHost, s));
    if (nfl > 0) F3D_HIP(hipMemcpyAsync(cfl, df, sizeof(double) * nf, hipMemcpyDeviceToHost, s));
    F3D_HIP(hipStreamSynchronize(s));
    return F3D_OK;
}

int f3d_restart_write(const char *path, int nx, int ny, int nz, int nc, int lonlat,
                      const double *ch, const double *cload, const f3d_run_info *info)
{
    if (!path || !ch || !cload || !info || nx < 1 || ny < 1 || nz < 1 || nc < 1) {
        set_error("f3d_restart_write: null argument or empty grid");
        return F3D_EINVAL;
    }
    const size_t nG = (size_t)(nx + 2) * (ny + 2) * (nz + 2) * nc, nL = (size_t)nx * ny * (nc + 1);
    if (nG * 4 + nL * 4 > ((uint64_t)1 << 40)) {
        set_error("f3d_restart_write: state too large for one file");
        return F3D_EINVAL;
    }
    // header: the dimensions, attributes and variables of wrirst.f90:26-145
    // in its order (the Fortran API lists a variable's dimensions fastest
    // first, the file slowest first)
    NcOut o;
    o.b = {'C', 'D', 'F', 2};
    o.u32(0);
    const std::pair<const char *, uint32_t> dims[] = {
        {"nx", (uint32_t)nx},      {"ny", (uint32_t)ny},      {"nc", (uint32_t)nc},
        {"nc1", (uint32_t)nc + 1}, {"nx2", (uint32_t)nx + 2}, {"ny2", (uint32_t)ny + 2},
        {"nz2", (uint32_t)nz + 2}};
    o.u32(NC_DIMENSION);
    o.u32(7);
    for (const auto &d : dims) {
        o.name(d.first);
        o.u32(d.second);
    }
    const bool ll = lonlat != 0;
    const std::vector<NcAtt> gatt = {
        {"TITLE", NC_CHAR, "Fall3d 6.2 results", {}},
        {"COORDINATES", NC_CHAR, ll ? "LON-LAT" : "UTM", {}},
        {"YEAR", NC_INT, "", {(double)info->year}},
        {"MONTH", NC_INT, "", {(double)info->month}},
        {"DAY", NC_INT, "", {(double)info->day}},
        {"RUN_START", NC_INT, "", {(double)info->run_start}},
        {"RUN_END", NC_INT, "", {(double)info->run_end}},
        {"ERUPTED_MASS", NC_DOUBLE, "", {info->erumass}},
        {"OUT_MASS", NC_DOUBLE, "", {info->outmass}},
        {ll ? "LONMIN" : "XMIN", NC_DOUBLE, "", {info->xmin}},
        {ll ? "LATMIN" : "YMIN", NC_DOUBLE, "", {info->ymin}},
        {ll ? "LONMAX" : "XMAX", NC_DOUBLE, "", {info->xmax}},
        {ll ? "LATMAX" : "YMAX", NC_DOUBLE, "", {info->ymax}}};
    put_atts(o, gatt);
    o.u32(NC_VARIABLE);
    o.u32(2);
    const uint64_t vsize_c = nG * 4, vsize_l = nL * 4;
    size_t begin_at[2];
    // CON (nx2,ny2,nz2,nc) and LOAD (nx,ny,nc1)
    o.name("CON");
    o.u32(4);
    for (uint32_t d : {2u, 6u, 5u, 4u}) o.u32(d);
    put_atts(o, {{"units", NC_CHAR, "gr/m3", {}}, {"description", NC_CHAR, "Scaled concentration", {}}});
    o.u32(NC_FLOAT);
    o.u32(vsize_c < 0xFFFFFFFFull ? (uint32_t)vsize_c : 0xFFFFFFFFu);
    begin_at[0] = o.b.size();
    o.u64(0);
    o.name("LOAD");
    o.u32(3);
    for (uint32_t d : {3u, 1u, 0u}) o.u32(d);
    put_atts(o, {{"units", NC_CHAR, "kg/m2", {}}, {"description", NC_CHAR, "Scaled ground deposit load", {}}});
    o.u32(NC_FLOAT);
    o.u32(vsize_l < 0xFFFFFFFFull ? (uint32_t)vsize_l : 0xFFFFFFFFu);
    begin_at[1] = o.b.size();
    o.u64(0);
    const uint64_t b0 = o.b.size(), b1 = b0 + (vsize_c + 3) / 4 * 4;
    for (int k = 0; k < 2; ++k) {
        const uint64_t v = k == 0 ? b0 : b1;
        for (int sft = 56, q = 0; sft >= 0; sft -= 8, ++q) o.b[begin_at[k] + q] = (unsigned char)(v >> sft);
    }
    FILE *f = std::fopen(path, "wb");
    if (!f) {
        set_error("f3d_restart_write: cannot open %s", path);
        return F3D_EINVAL;
    }
    bool ok = std::fwrite(o.b.data(), 1, o.b.size(), f) == o.b.size();
    // data: 32-bit floats (rounded to nearest, as NF90_FLOAT stores a
    // real(rp) array), big-endian, in the Fortran order of
    // c(0:nx+1,0:ny+1,0:nz+1,nc) and cload(nx,ny,0:nc)
    std::vector<unsigned char> chunk;
    const size_t kChunk = (size_t)1 << 20;
    auto put_floats = [&](const double *src, size_t n) {
        chunk.resize(4 * std::min(n, kChunk));
        for (size_t i0 = 0; i0 < n && ok; i0 += kChunk) {
            const size_t m = std::min(n - i0, kChunk);
            for (size_t i = 0; i < m; ++i) {
                const float v = (float)src[i0 + i];
                uint32_t u;
                std::memcpy(&u, &v, 4);
                chunk[4 * i] = (unsigned char)(u >> 24);
                chunk[4 * i + 1] = (unsigned char)(u >> 16);
                chunk[4 * i + 2] = (unsigned char)(u >> 8);
                chunk[4 * i + 3] = (unsigned char)u;
            }
            ok = ok && std::fwrite(chunk.data(), 1, 4 * m, f) == 4 * m;
        }
    };
    put_floats(ch, nG);
    put_floats(cload, nL);
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) {
        set_error("f3d_restart_write: write to %s failed", path);
        return F3D_EINVAL;
    }
    return F3D_OK;
}

int f3d_restart_read(const char *path, int nx, int ny, int nz, int nc, double *ch,
                     double *cload, f3d_run_info *info)
{
    if (!path || nx < 1 || ny < 1 || nz < 1 || nc < 1) {
        set_error("f3d_restart_read: null path or empty grid");
        return F3D_EINVAL;
    }
    FILE *f = std::fopen(path, "rb");
    if (!f) {
        set_error("f3d_restart_read: cannot open %s", path);
        return F3D_EINVAL;
    }
    std::fseek(f, 0, SEEK_END);
    const long fsz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    std::vector<unsigned char> head((size_t)std::min<long>(std::max<long>(fsz, 0), 1 << 20));
    const bool got = std::fread(head.data(), 1, head.size(), f) == head.size();
    NcIn in{head.data(), head.size()};
    std::vector<std::pair<std::string, uint64_t>> dims;
    std::vector<std::pair<std::string, double>> gatt;
    std::vector<NcVarIn> vars;
    if (!got || !parse_header(in, dims, gatt, vars)) {
        std::fclose(f);
        set_error("f3d_restart_read: %s is not a netCDF classic file", path);
        return F3D_EINVAL;
    }
    // rearst.f90:34-40,141-144: the grid of the file must be the caller's
    long nx_r = -1, ny_r = -1, nz_r = -1, nc_r = -1;
    for (const auto &d : dims) {
        if (d.first == "nx") nx_r = (long)d.second;
        if (d.first == "ny") ny_r = (long)d.second;
        if (d.first == "nz2") nz_r = (long)d.second - 2;
        if (d.first == "nc") nc_r = (long)d.second;
    }
    if (nx_r != nx || ny_r != ny || nz_r != nz || nc_r != nc) {
        std::fclose(f);
        set_error("f3d_restart_read: restart inconsistency: file grid %ldx%ldx%ld, %ld classes; "
                  "expected %dx%dx%d, %d classes", nx_r, ny_r, nz_r, nc_r, nx, ny, nz, nc);
        return F3D_EINVAL;
    }
    const NcVarIn *vc = nullptr, *vl = nullptr;
    for (const NcVarIn &v : vars) {
        if (v.name == "CON") vc = &v;
        if (v.name == "LOAD") vl = &v;
    }
    const size_t nG = (size_t)(nx + 2) * (ny + 2) * (nz + 2) * nc, nL = (size_t)nx * ny * (nc + 1);
    auto shape_is = [&](const NcVarIn *v, std::initializer_list<uint64_t> want) {
        if (!v || v->type != NC_FLOAT || v->dims.size() != want.size()) return false;
        size_t k = 0;
        for (uint64_t w : want) {
            if (v->dims[k] >= dims.size() || dims[v->dims[k]].second != w) return false;
            ++k;
        }
        return true;
    };
    if (!shape_is(vc, {(uint64_t)nc, (uint64_t)nz + 2, (uint64_t)ny + 2, (uint64_t)nx + 2}) ||
        !shape_is(vl, {(uint64_t)nc + 1, (uint64_t)ny, (uint64_t)nx}) ||
        vc->begin + 4 * nG > (uint64_t)fsz || vl->begin + 4 * nL > (uint64_t)fsz) {
        std::fclose(f);
        set_error("f3d_restart_read: %s has no float CON(nx2,ny2,nz2,nc) and LOAD(nx,ny,nc1)", path);
        return F3D_EINVAL;
    }
    auto read_floats = [&](uint64_t begin, double *dst, size_t n) {
        if (!dst) return true;
        std::vector<unsigned char> raw(4 * n);
        if (std::fseek(f, (long)begin, SEEK_SET) != 0 || std::fread(raw.data(), 1, raw.size(), f) != raw.size())
            return false;
        for (size_t i = 0; i < n; ++i) {
            const uint32_t u = ((uint32_t)raw[4 * i] << 24) | ((uint32_t)raw[4 * i + 1] << 16) |
                               ((uint32_t)raw[4 * i + 2] << 8) | raw[4 * i + 3];
            float v;
            std::memcpy(&v, &u, 4);
            dst[i] = (double)v;
        }
        return true;
    };
    const bool ok = read_floats(vc->begin, ch, nG) && read_floats(vl->begin, cload, nL);
    std::fclose(f);
    if (!ok) {
        set_error("f3d_restart_read: %s is truncated", path);
        return F3D_EINVAL;
    }
    if (info) {
        auto att = [&](const char *name, double dflt) {
            for (const auto &a : gatt)
                if (a.first == name) return a.second;
            return dflt;
        };
        bool ll = false;
        for (const auto &a : gatt) ll = ll || a.first == "LONMIN";
        info->year = (int)att("YEAR", 0);
        info->month = (int)att("MONTH", 0);
        info->day = (int)att("DAY", 0);
        info->run_start = (int)att("RUN_START", 0);
        info->run_end = (int)att("RUN_END", 0);
        info->xmin = att(ll ? "LONMIN" : "XMIN", 0.0);
        info->ymin = att(ll ? "LATMIN" : "YMIN", 0.0);
        info->xmax = att(ll ? "LONMAX" : "XMAX", 0.0);
        info->ymax = att(ll ? "LATMAX" : "YMAX", 0.0);
        info->erumass = att("ERUPTED_MASS", 0.0);
        info->outmass = att("OUT_MASS", 0.0);
    }
    return F3D_OK;
}

int f3d_write_restart(f3d_ctx *c, int member, const char *path, const f3d_run_info *info)
{
    if (!check_ctx(c)) return F3D_EINVAL;
    const Grid &g = c->g;
    if (!path || !info) {
        set_error("f3d_write_restart: null path or run info");
        return F3D_EINVAL;
    }
    if (member < 0 || member >= g.nm) {
        set_error("f3d_write_restart: member %d out of range (ensemble of %d)", member, g.nm);
        return F3D_EINVAL;
    }
    // the state, the deposit and the clock of the member
    const size_t nG = (size_t)g.G * g.nc, nL = (size_t)g.NXY * (g.nc + 1);
    std::vector<double> ch(nG), cdry(nL), cwet(nL), load(nL);
    TRY(f3d_get_member_state(c, member, ch.data(), cdry.data()));
    hipStream_t s = c->stream;
    F3D_HIP(hipMemcpyAsync(cwet.data(), c->d_cwet.as<double>() + nL * member, sizeof(double) * nL,
                           hipMemcpyDeviceToHost, s));
    double outmass = 0.0;
    F3D_HIP(hipMemcpyAsync(&outmass, c->d_outmass.as<double>() + member, sizeof(double),
                           hipMemcpyDeviceToHost, s));
    F3D_HIP(hipStreamSynchronize(s));
    // getc0 (endstep.f90:80-106), then cload = cdry + cwet (endstep.f90:35)
    for (long t = 0; t < g.NXY; ++t) {
        double d0 = 0.0, w0 = 0.0;
        for (int q = 1; q <= g.nc; ++q) {
            d0 = d0 + cdry[q * g.NXY + t];
            w0 = w0 + cwet[q * g.NXY + t];
        }
        load[t] = d0 + w0;
        for (int q = 1; q <= g.nc; ++q) load[q * g.NXY + t] = cdry[q * g.NXY + t] + cwet[q * g.NXY + t];
    }
    f3d_run_info run = *info;
    run.erumass = c->erumass[member];
    run.outmass = outmass;
    return f3d_restart_write(path, g.nx, g.ny, g.nz, g.nc, c->lonlat, ch.data(), load.data(), &run);
}

int f3d_read_restart(f3d_ctx *c, const char *path, double *cload, f3d_run_info *info)
{
    if (!check_ctx(c)) return F3D_EINVAL;
    const Grid &g = c->g;
    if (g.nm != 1) {
        set_error("f3d_read_restart: single-member contexts only");
        return F3D_EINVAL;
    }
    const size_t nG = (size_t)g.G * g.nc, nL = (size_t)g.NXY * (g.nc + 1);
    std::vector<double> ch(nG), zero(nL, 0.0);
    f3d_run_info run{};
    TRY(f3d_restart_read(path, g.nx, g.ny, g.nz, g.nc, ch.data(), cload, &run));
    // rearst.f90:186-187: c (and cload) from the file; the reference's cdry
    // and cwet start again from zero
    TRY(f3d_set_state(c, ch.data(), zero.data()));
    F3D_HIP(hipMemcpyAsync(c->d_cwet.p, zero.data(), sizeof(double) * nL, hipMemcpyHostToDevice,
                           c->stream));
    F3D_HIP(hipStreamSynchronize(c->stream));
    // rearst.f90:98-109: ERUPTED_MASS and OUT_MASS
    TRY(f3d_set_clock(c, c->time, c->iiter, run.erumass, run.outmass));
    if (info) *info = run;
    return F3D_OK;
}

}  // extern "C"
