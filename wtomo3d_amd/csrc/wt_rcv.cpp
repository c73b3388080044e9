// Receiver extraction (makeuh.f:1-172 with srcsetup.f / srcreg.f): every
// receiver output is a fixed linear functional of the solution, so the
// stations are compiled once into a sparse extraction matrix E (rows:
// u, v, w per receiver, then one pressure per hydrophone; columns: unknown
// indices) and applied to whole solution blocks on the GPU.
#include <cmath>
#include <complex>
#include <vector>

#include "wt_rhs.h"
#include "wt_rt.h"

namespace {

typedef std::complex<double> zc;

// Kaiser window parameter per sinc half-length (Hicks 2002, Geophysics 67,
// accuracy to half the Nyquist wavenumber); half-lengths above 9 use 14.18.
double kaiser_b(int r)
{
    static const double b[10] = {0.0, 1.24, 2.94, 4.53, 6.31, 7.91, 9.42, 10.95, 12.53, 14.09};
    return (r >= 1 && r <= 9) ? b[r] : 14.18;
}

// I0 by the polynomial approximations of Abramowitz & Stegun 9.8.1-9.8.2
double bessel_i0(double x)
{
    const double ax = std::fabs(x);
    if (ax < 3.75) {
        const double t = (x / 3.75) * (x / 3.75);
        return 1.0 + t * (3.5156229 + t * (3.0899424 + t * (1.2067492 + t * (0.2659732 + t * (0.0360768 +
                                                                                              t * 0.0045813)))));
    }
    const double t = 3.75 / ax;
    const double p = 0.39894228 + t * (0.01328592 + t * (0.00225319 + t * (-0.00157565 + t * (0.00916281 + t * (
                                        -0.02057706 + t * (0.02635537 + t * (-0.01647633 + t * 0.00392377)))))));
    return std::exp(ax) / std::sqrt(ax) * p;
}

double kaiser_sinc(int r, double x)
{
    const double s = x <= 1e-10 ? 1.0 : std::sin(M_PI * x) / (M_PI * x);
    const double lx = x / (double)r;
    if (!(lx < 1.0)) return 0.0;
    const double b = kaiser_b(r);
    return s * bessel_i0(b * std::sqrt(1.0 - lx * lx)) / bessel_i0(b);
}

// discrete Dirac pulse of half-width h at distance x (Alterman & Aboudi 1970)
double dirac(double h, double x)
{
    if (x < h) return (1.0 - x * x / (2.0 * h * h)) / (2.0 * h);
    if (x < 2.0 * h) return (1.0 + (x * x - 4.0 * x * h) / (4.0 * h * h)) / h;
    return 0.0;
}

struct Grid {
    int nx, ny, nz;
    double dx;
    bool fs[6];  // free surface: top, +x, bottom, -x, +y, -y
};

struct Tap {
    int64_t node;  // 0-based node index (z fastest, then x, then y)
    double w;
};

// Taps of the value at (x, y, z): nearest node for ireg = 0, else the region
// weights with mirrored (sign-flipped) taps beyond free surfaces and dropped
// taps beyond other boundaries.  Returns WT_ERR_ARG when a tap would fall
// outside the grid.
int station_taps(const Grid& g, double x, double y, double z, double spread, int ireg, std::vector<Tap>& out)
{
    out.clear();
    auto node = [&](int iz, int ix, int iy) { return ((int64_t)(iy - 1) * g.nx + (ix - 1)) * g.nz + (iz - 1); };
    if (ireg == 0) {
        const int ix = (int)(0.5 + x / g.dx) + 1, iy = (int)(0.5 + y / g.dx) + 1, iz = (int)(0.5 + z / g.dx) + 1;
        if (ix < 1 || ix > g.nx || iy < 1 || iy > g.ny || iz < 1 || iz > g.nz) return WT_ERR_ARG;
        out.push_back({node(iz, ix, iy), 1.0});
        return WT_OK;
    }
    const int ixs = (int)(x / g.dx) + 1, iys = (int)(y / g.dx) + 1, izs = (int)(z / g.dx) + 1;
    const double ox = x - (double)(ixs - 1) * g.dx, oy = y - (double)(iys - 1) * g.dx, oz = z - (double)(izs - 1) * g.dx;
    const bool point = spread <= 0.5;
    const int h = point ? ireg : (int)(2 * spread);
    const int m = point ? 2 * ireg + 1 : (int)(4 * spread) + 1;
    const double c = point ? ireg + 1 : (int)(2 * spread) + 1;
    // separable weights along each axis, region index 1..m
    std::vector<double> wx(m + 1), wy(m + 1), wz(m + 1);
    for (int i = 1; i <= m; ++i) {
        const double dxr = std::fabs(i - (c + ox / g.dx)), dyr = std::fabs(i - (c + oy / g.dx)),
                     dzr = std::fabs(i - (c + oz / g.dx));
        wx[i] = point ? kaiser_sinc(ireg, dxr) : dirac(spread, dxr);
        wy[i] = point ? kaiser_sinc(ireg, dyr) : dirac(spread, dyr);
        wz[i] = point ? kaiser_sinc(ireg, dzr) : dirac(spread, dzr);
    }
    double norm = 1.0;
    if (!point) {
        // the distributed weights are normalised over the whole m^3 block
        double sx = 0, sy = 0, sz = 0;
        for (int i = 1; i <= m; ++i) { sx += wx[i]; sy += wy[i]; sz += wz[i]; }
        norm = sx * sy * sz;
    }
    const int zlo = izs - h, zhi = izs + h, xlo = ixs - h, xhi = ixs + h, ylo = iys - h, yhi = iys + h;
    const int zmin = g.fs[0] ? zlo : std::max(zlo, 1), zmax = g.fs[2] ? zhi : std::min(zhi, g.nz);
    const int xmin = g.fs[3] ? xlo : std::max(xlo, 1), xmax = g.fs[1] ? xhi : std::min(xhi, g.nx);
    const int ymin = g.fs[5] ? ylo : std::max(ylo, 1), ymax = g.fs[4] ? yhi : std::min(yhi, g.ny);
    auto mirror = [](int i, int n) { return i < 1 ? 2 - i : (i > n ? 2 * n - i : i); };
    for (int iz = zmin; iz <= zmax; ++iz) {
        const int izf = mirror(iz, g.nz);
        for (int ix = xmin; ix <= xmax; ++ix) {
            const int ixf = mirror(ix, g.nx);
            for (int iy = ymin; iy <= ymax; ++iy) {
                const int iyf = mirror(iy, g.ny);
                if (izf < 1 || izf > g.nz || ixf < 1 || ixf > g.nx || iyf < 1 || iyf > g.ny) return WT_ERR_ARG;
                const double w = wz[iz - zlo + 1] * wy[iy - ylo + 1] * wx[ix - xlo + 1] / norm;
                const bool inside = iz == izf && ix == ixf && iy == iyf;
                out.push_back({node(izf, ixf, iyf), inside ? w : -w});
            }
        }
    }
    return WT_OK;
}

}  // namespace

struct wt_receivers {
    int rows = 0;
    int32_t* rowptr = nullptr;
    int32_t* col = nullptr;
    cpx<double>* val = nullptr;
};

extern "C" {

int wt_receivers_create(int nx, int ny, int nz, double dx, const int32_t* freesurf, int ng, const double* xg,
                        const double* yg, const double* zg, double gspread, int igreg, const double* gfac, int nr,
                        const double* xr, const double* yr, const double* zr, double rspread, int irreg,
                        const void* lam, const void* mu, double omega_re, double omega_im, wt_receivers** out)
{
    if (!out || nx < 1 || ny < 1 || nz < 1 || !(dx > 0) || ng < 0 || nr < 0 || ng + nr < 1) return WT_ERR_ARG;
    if ((ng && (!xg || !yg || !zg)) || (nr && (!xr || !yr || !zr || !lam || !mu))) return WT_ERR_ARG;
    // the reference's region array is 9 nodes wide (srcreg.f: fs(9,9,9))
    auto valid = [](double spread, int ireg) {
        if (ireg < 0 || ireg > 4) return false;
        if (ireg > 0 && spread > 0.5 && (int)(4 * spread) + 1 > 9) return false;
        return true;
    };
    if ((ng && !valid(gspread, igreg)) || (nr && !valid(rspread, irreg))) return WT_ERR_ARG;
    Grid g{nx, ny, nz, dx, {false, false, false, false, false, false}};
    for (int k = 0; k < 6 && freesurf; ++k) g.fs[k] = freesurf[k] != 0;
    std::vector<int32_t> rp(1, 0);
    std::vector<int32_t> cols;
    std::vector<zc> vals;
    std::vector<Tap> taps;
    auto add = [&](int comp, zc f) {
        for (const Tap& t : taps) {
            cols.push_back((int32_t)(3 * t.node + comp));
            vals.push_back(f * t.w);
        }
    };
    for (int k = 0; k < ng; ++k) {
        const zc f = gfac ? zc(gfac[2 * k], gfac[2 * k + 1]) : zc(1.0, 0.0);
        int rc = station_taps(g, xg[k], yg[k], zg[k], gspread, igreg, taps);
        if (rc) return rc;
        for (int comp = 0; comp < 3; ++comp) {
            add(comp, f);
            rp.push_back((int32_t)cols.size());
        }
    }
    // hydrophones: (lambda + 2 mu / 3) at the nearest node times the
    // divergence from central differences over 2 d (-i omega)
    const zc cfact = zc(0.0, -1.0) * zc(omega_re, omega_im);
    const cpx<double>* L = (const cpx<double>*)lam;
    const cpx<double>* M = (const cpx<double>*)mu;
    for (int k = 0; k < nr; ++k) {
        const int ix = std::min(std::max((int)(1.5 + xr[k] / dx), 1), nx);
        const int iy = std::min(std::max((int)(1.5 + yr[k] / dx), 1), ny);
        const int iz = std::min(std::max((int)(1.5 + zr[k] / dx), 1), nz);
        const int64_t nd = ((int64_t)(iy - 1) * nx + (ix - 1)) * nz + (iz - 1);
        const zc modulus = zc(L[nd].re, L[nd].im) + 2.0 * zc(M[nd].re, M[nd].im) / 3.0;
        const zc c = -modulus / (2.0 * dx * cfact);
        const double sh[3][3] = {{dx, 0, 0}, {0, dx, 0}, {0, 0, dx}};
        for (int comp = 0; comp < 3; ++comp) {
            int rc = station_taps(g, xr[k] + sh[comp][0], yr[k] + sh[comp][1], zr[k] + sh[comp][2], rspread, irreg,
                                  taps);
            if (rc) return rc;
            add(comp, c);
            rc = station_taps(g, xr[k] - sh[comp][0], yr[k] - sh[comp][1], zr[k] - sh[comp][2], rspread, irreg, taps);
            if (rc) return rc;
            add(comp, -c);
        }
        rp.push_back((int32_t)cols.size());
    }
    wt_receivers* r = new (std::nothrow) wt_receivers;
    if (!r) return WT_ERR_NOMEM;
    r->rows = (int)rp.size() - 1;
    const size_t ne = cols.size() ? cols.size() : 1;
    if (rt_malloc(&r->rowptr, rp.size() * 4) != hipSuccess || rt_malloc(&r->col, ne * 4) != hipSuccess ||
        rt_malloc(&r->val, ne * 16) != hipSuccess) {
        (void)rt_free(r->rowptr);
        (void)rt_free(r->col);
        (void)rt_free(r->val);
        delete r;
        return WT_ERR_NOMEM;
    }
    std::vector<cpx<double>> v(vals.size());
    for (size_t i = 0; i < vals.size(); ++i) v[i] = cpx<double>(vals[i].real(), vals[i].imag());
    if (rt_memcpy(r->rowptr, rp.data(), rp.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
        (cols.size() && rt_memcpy(r->col, cols.data(), cols.size() * 4, hipMemcpyHostToDevice) != hipSuccess) ||
        (v.size() && rt_memcpy(r->val, v.data(), v.size() * 16, hipMemcpyHostToDevice) != hipSuccess)) {
        (void)rt_free(r->rowptr);
        (void)rt_free(r->col);
        (void)rt_free(r->val);
        delete r;
        return WT_ERR_HIP;
    }
    *out = r;
    return WT_OK;
}

int wt_receivers_rows(const wt_receivers* r) { return r ? r->rows : 0; }

int wt_receivers_extract(const wt_receivers* r, int nrhs, const void* d_field, void* d_out, void* stream)
{
    if (!r || !d_field || !d_out || nrhs < 1) return WT_ERR_ARG;
    const hipStream_t st = (hipStream_t)stream;
    wt_rt_stream_guard g(stream);
    int rc = wt_launch_rows(r->rowptr, r->col, r->val, (const cpx<double>*)d_field, (cpx<double>*)d_out, r->rows,
                            nrhs, st);
    if (rc) return rc;
    WT_HIP(hipStreamSynchronize(st));
    return WT_OK;
}

int wt_receivers_destroy(wt_receivers* r)
{
    if (!r) return WT_ERR_ARG;
    (void)rt_free(r->rowptr);
    (void)rt_free(r->col);
    (void)rt_free(r->val);
    delete r;
    return WT_OK;
}

}  // extern "C"
