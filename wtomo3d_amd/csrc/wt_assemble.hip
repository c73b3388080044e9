// Elastic 27-point stencil + PML assembly into CSR on MI355X (gfx950).
//
// Replaces ASMBLE (asmble.f90:96-601) and the per-node physics it calls:
// SETCOEFS3D (setcoefs3d.f:79-233), CALCWTS (calcwts.f:28-71), the region
// dispatch of FDFD3D (fdfd3d.f:50-115), PFDFD3D (pfdfd3d.f:133-576),
// PML3D (pml3d.f:167-800) and FIXED3D (fixed3d.f:41-343).  The CSR pattern
// kernel reproduces SREL3D (srgraphs.f:638-734) followed by SPARSE_ADJ2CRS
// (sparse.f:13-49) from closed forms.
//
// Design: a block of 256 threads owns a segment of 64 z-nodes of one (x, y)
// column.  The clamped 3x3 column neighbourhood of lambda, mu, rho is staged
// in LDS, the octant-cell moduli are formed once per z-interface into an LDS
// table, and lane l of every wave forms node l's mass weights and PML
// stretch factors in registers.  A node's output is 3 rows of up to 81
// entries (neighbours ascending in y, x, z), and the rows of consecutive
// nodes follow each other in the CSR arrays, so a chunk of 16 nodes is one
// contiguous span.  The block builds each chunk's span in LDS -- each wave
// takes its share of the 27 neighbour labels, the label uniform over 16 lanes
// (the eight corners and the three families of four edges four at a time and
// the three pairs of faces two at a time, one formula with per-lane cells and
// signs; the centre alone) -- and writes it front to back, consecutive
// threads on consecutive entries, while the next chunk is evaluated.  The
// interior stencil is the stretched form with r1 = 1, r2 = 0 (identical
// values; the reference notes the identity at fdfd3d.f:56-57).  CSR offsets
// follow from closed forms, so no atomics or scans.
//
// Measured at 96x96x64 in FP64: one wave per segment storing each neighbour
// column's 144-byte row runs took 1.05 ms; the contiguous spans with the
// per-interface cell table 0.69 ms, with the four-wide label groups 0.62 ms,
// with the coefficients formed by one wave 0.61 ms, with the face pairs and
// the waves' label sets balanced by measurement 0.59 ms; the span stores
// alone (no values computed) take 0.37 ms.  The labels cost 0.15 ms of the
// rest, the tile loads and barriers the remainder, with occupancy capped at
// two blocks per CU by the FP64 registers (256) and the 76 KB of LDS.
#include "wt_common.h"

#include <type_traits>

#pragma clang fp contract(off)

namespace {

constexpr int ASM_WAVE = 64;

// Number of existing neighbours along one axis and prefix sums of it.
__host__ __device__ inline int64_t axis_count(int i, int n) { return 1 + (i > 1) + (i < n); }
__host__ __device__ inline int64_t axis_prefix(int i, int n)
{
    // sum_{i' < i} axis_count(i', n), 1 <= i <= n + 1
    int64_t a = i - 1, b = i - 2 > 0 ? i - 2 : 0, c = (i - 1 < n - 1) ? i - 1 : n - 1;
    return a + b + c;
}

// Offset (0-based, in nonzeros) of the first row of node (iz, ix, iy).
__host__ __device__ inline int64_t node_offset(int iz, int ix, int iy, int nx, int ny, int nz)
{
    int64_t Sx = axis_prefix(nx + 1, nx), Sz = axis_prefix(nz + 1, nz);
    int64_t cy = axis_count(iy, ny), cx = axis_count(ix, nx);
    return 9 * (axis_prefix(iy, ny) * Sx * Sz + cy * axis_prefix(ix, nx) * Sz +
                cy * cx * axis_prefix(iz, nz));
}

template <typename R>
struct NodeCoef {
    cpx<R> rx2, ry2, rz2, rxy, rxz, ryz;  // r1-derived scalings
    cpx<R> rx1, ry1, rz1;                // r2 / (2 dx) / 4
    R wm2o6, wm3o12, wm4o8;
    cpx<R> mc;                           // centre mass term (complex)
    // Octant-cell mu and lambda, index 1..8, held in LDS: cells 1..4 lie
    // above the node (cup: mu 1..4, then lambda 1..4), cells 5..8 below it
    // (cdn: mu 5..8, then lambda 5..8).
    // (The numeric members come first: launder() walks them as an R array.)
    const cpx<R>* cup;
    const cpx<R>* cdn;
    __device__ cpx<R> U(int i) const { return i <= 4 ? cup[i - 1] : cdn[i - 5]; }
    __device__ cpx<R> D(int i) const { return i <= 4 ? cup[3 + i] : cdn[i - 1]; }
};

template <typename R>
__device__ inline R csqrt_re(cpx<R> z)
{
    if (z.im == R(0)) return z.re >= R(0) ? sqrt(z.re) : R(0);
    R m = hypot(z.re, z.im);
    return sqrt((m + z.re) / R(2));
}

// Stretch factors along one axis (pml3d.f:207-295 with kmax = 1, fdom = 0).
template <typename R>
__device__ inline bool pml_axis(int i, int n, int ipml, R pmld, R dx, R pmlf, R vpreal,
                                cpx<R> eyeom, cpx<R>& r1, cpx<R>& r2)
{
    R dpml = 0;
    int isn = 0;
    if (i <= ipml) { dpml = pmld - R(i - 1) * dx; isn = 1; }
    else if (n - i + 1 <= ipml) { dpml = pmld - R(n - i) * dx; isn = -1; }
    if (dpml < 0) return false;
    R dn = dpml * dpml;
    dn = dn * vpreal;
    dn = dn * pmlf;
    r1 = cdiv(eyeom, cpx<R>(eyeom.re + dn, eyeom.im));
    cpx<R> r1sq = r1 * r1;
    cpx<R> t = cdiv(eyeom, eyeom * eyeom);
    cpx<R> q = (R(isn) * R(2)) * r1sq;
    q = q * r1;
    q = dpml * q;
    r2 = q * ((vpreal * pmlf) * t);
    return true;
}

// The 3x3 block (row-major u,v,w x u,v,w) coupling a node to neighbour k,
// k = 9(oy+1) + 3(ox+1) + (oz+1).  ro is the density at that neighbour as the
// reference's mass lumping sees it.
template <typename R>
__device__ __forceinline__ void label_block(int k, const NodeCoef<R>& c, R ro, cpx<R> v[9])
{
    auto lmu = [&](int i) { return c.D(i) - c.U(i); };
    auto lpu = [&](int i) { return c.D(i) + c.U(i); };
    auto lp2u = [&](int i) { return (c.D(i) + c.U(i)) + c.U(i); };
    const cpx<R> rx2 = c.rx2, ry2 = c.ry2, rz2 = c.rz2, rxy = c.rxy, rxz = c.rxz, ryz = c.ryz;
    const cpx<R> rx1 = c.rx1, ry1 = c.ry1, rz1 = c.rz1;
    // Diagonal of the 8 corner blocks: q1*L + qs*U +- ra*L +- rb*U + mass.
    auto corner = [&](cpx<R> q1, cpx<R> qs, int sa, cpx<R> ra, int sb, cpx<R> rb, int cell, cpx<R> mf) {
        cpx<R> L = lp2u(cell), U = c.U(cell);
        cpx<R> s = q1 * L + qs * U;
        cpx<R> ta = ra * L, tb = rb * U;
        s = s + (sa > 0 ? ta : -ta);
        s = s + (sb > 0 ? tb : -tb);
        return s + mf;
    };
    auto put = [&](cpx<R> uu, cpx<R> vv, cpx<R> ww, cpx<R> uv, cpx<R> vu, cpx<R> uw, cpx<R> wu,
                   cpx<R> vw, cpx<R> wv) {
        v[0] = uu; v[1] = uv; v[2] = uw;
        v[3] = vu; v[4] = vv; v[5] = vw;
        v[6] = wu; v[7] = wv; v[8] = ww;
    };
    const cpx<R> rxpy = rx2 + ry2, rxpz = rx2 + rz2, rypz = ry2 + rz2;
    const cpx<R> r2xpy = rx1 + ry1, r2xmy = rx1 - ry1, r2xpz = rx1 + rz1, r2xmz = rx1 - rz1;
    const cpx<R> r2ypz = ry1 + rz1, r2ymz = ry1 - rz1;
    const cpx<R> m4(c.wm4o8 * ro, R(0)), m3(c.wm3o12 * ro, R(0)), m2(c.wm2o6 * ro, R(0));
    cpx<R> t0, t1, t2, t3, t4, t5;
    switch (k) {
    // ---------------- oy = -1 ----------------
    case 0:  // a1m
        t0 = rxy * lpu(1); t1 = rxz * lpu(1); t2 = ryz * lpu(1);
        put(corner(rx2, rypz, -1, rx1, -1, r2ypz, 1, m4), corner(ry2, rxpz, -1, ry1, -1, r2xpz, 1, m4),
            corner(rz2, rxpy, -1, rz1, -1, r2xpy, 1, m4), t0, t0, t1, t1, t2, t2);
        break;
    case 1:  // a2m
        t0 = -(rz2 * (c.U(1) + c.U(5))) + m3; t1 = -(rz2 * (lp2u(1) + lp2u(5))) + m3;
        t2 = rxy * (lpu(1) + lpu(5)); t3 = -(rxz * (lmu(1) - lmu(5))); t4 = -(ryz * (lmu(1) - lmu(5)));
        put(t0, t0, t1, t2, t2, t3, -t3, t4, -t4);
        break;
    case 2:  // a3m
        t0 = rxy * lpu(5); t1 = -(rxz * lpu(5)); t2 = -(ryz * lpu(5));
        put(corner(rx2, rypz, -1, rx1, -1, r2ymz, 5, m4), corner(ry2, rxpz, -1, ry1, -1, r2xmz, 5, m4),
            corner(rz2, rxpy, +1, rz1, -1, r2xmy, 5, m4), t0, t0, t1, t1, t2, t2);
        break;
    case 3:  // a4m
        t0 = -(rx2 * (lp2u(1) + lp2u(3))) + m3; t1 = -(rx2 * (c.U(1) + c.U(3))) + m3;
        t2 = rxy * (lmu(1) - lmu(3)); t3 = rxz * (lmu(1) - lmu(3)); t4 = ryz * (lpu(1) + lpu(3));
        put(t0, t0, t1, t2, -t2, t3, -t3, t4, t4);
        break;
    case 4:  // a5m
        t0 = ry2 * (((c.U(1) + c.U(7)) + c.U(5)) + c.U(3)) + m2;
        t2 = rxy * (((lmu(1) - lmu(7)) + lmu(5)) - lmu(3));
        t3 = -(rxz * (((lpu(1) + lpu(7)) - lpu(5)) - lpu(3)));
        t4 = -(ryz * (((lmu(1) - lmu(7)) - lmu(5)) + lmu(3)));
        put(t0, t0, t0, t2, -t2, t3, t3, t4, -t4);
        break;
    case 5:  // a6m
        t0 = -(rx2 * (lp2u(7) + lp2u(5))) + m3; t1 = -(rx2 * (c.U(7) + c.U(5))) + m3;
        t2 = rxy * (lmu(5) - lmu(7)); t3 = rxz * (lmu(7) - lmu(5)); t4 = -(ryz * (lpu(7) + lpu(5)));
        put(t0, t0, t1, t2, -t2, t3, -t3, t4, t4);
        break;
    case 6:  // a7m
        t0 = -(rxy * lpu(3)); t1 = -(rxz * lpu(3)); t2 = ryz * lpu(3);
        put(corner(rx2, rypz, +1, rx1, -1, r2ypz, 3, m4), corner(ry2, rxpz, -1, ry1, +1, r2xmz, 3, m4),
            corner(rz2, rxpy, -1, rz1, +1, r2xpy, 3, m4), t0, t0, t1, t1, t2, t2);
        break;
    case 7:  // a8m
        t0 = -(rz2 * (c.U(7) + c.U(3))) + m3; t1 = -(rz2 * (lp2u(7) + lp2u(3))) + m3;
        t2 = -(rxy * (lpu(7) + lpu(3))); t3 = -(rxz * (lmu(7) - lmu(3))); t4 = -(ryz * (lmu(3) - lmu(7)));
        put(t0, t0, t1, t2, t2, t3, -t3, t4, -t4);
        break;
    case 8:  // a9m
        t0 = -(rxy * lpu(7)); t1 = rxz * lpu(7); t2 = -(ryz * lpu(7));
        put(corner(rx2, rypz, +1, rx1, -1, r2ymz, 7, m4), corner(ry2, rxpz, -1, ry1, +1, r2xpz, 7, m4),
            corner(rz2, rxpy, +1, rz1, +1, r2xmy, 7, m4), t0, t0, t1, t1, t2, t2);
        break;
    // ---------------- oy = 0 ----------------
    case 9:  // a1n
        t0 = -(ry2 * (c.U(1) + c.U(2))) + m3;
        t2 = rxy * (lmu(2) - lmu(1)); t3 = rxz * (lpu(1) + lpu(2)); t4 = ryz * (lmu(1) - lmu(2));
        put(t0, t0, t0, t2, -t2, t3, t3, t4, -t4);
        break;
    case 10:  // a2n
        t0 = rx2 * (((lp2u(1) + lp2u(6)) + lp2u(2)) + lp2u(5)) + m2;
        t1 = rx2 * (((c.U(1) + c.U(6)) + c.U(2)) + c.U(5)) + m2;
        t2 = rxy * (((lmu(2) - lmu(5)) + lmu(6)) - lmu(1));
        t3 = -(rxz * (((lmu(2) - lmu(5)) + lmu(1)) - lmu(6)));
        t4 = ryz * (((lpu(2) + lpu(5)) - lpu(1)) - lpu(6));
        put(t0, t0, t1, t2, -t2, t3, -t3, t4, t4);
        break;
    case 11:  // a3n
        t0 = -(ry2 * (c.U(5) + c.U(6))) + m3;
        t2 = -(rxy * (lmu(5) - lmu(6))); t3 = -(rxz * (lpu(5) + lpu(6))); t4 = -(ryz * (lmu(5) - lmu(6)));
        put(t0, t0, t0, t2, -t2, t3, t3, t4, -t4);
        break;
    case 12:  // a4n
        t0 = rz2 * (((c.U(1) + c.U(4)) + c.U(2)) + c.U(3)) + m2;
        t1 = rz2 * (((lp2u(1) + lp2u(4)) + lp2u(2)) + lp2u(3)) + m2;
        t2 = -(rxy * (((lpu(1) + lpu(4)) - lpu(2)) - lpu(3)));
        t3 = rxz * (((lmu(1) - lmu(4)) + lmu(2)) - lmu(3));
        t4 = ryz * (((lmu(1) - lmu(4)) + lmu(3)) - lmu(2));
        put(t0, t0, t1, t2, t2, t3, -t3, t4, -t4);
        break;
    case 13: {  // a5n, the node itself
        cpx<R> ls = lp2u(1), us = c.U(1);
#pragma unroll
        for (int i = 2; i <= 8; ++i) { ls = ls + lp2u(i); us = us + c.U(i); }
        cpx<R> ex = ((((((lp2u(1) - lp2u(8)) + lp2u(2)) - lp2u(7)) + lp2u(5)) - lp2u(4)) + lp2u(6)) - lp2u(3);
        cpx<R> fa = ((c.U(1) - c.U(8)) + c.U(3)) - c.U(6);
        cpx<R> fb = ((c.U(7) - c.U(2)) + c.U(5)) - c.U(4);
        t0 = ((((-(rx2 * ls) - rypz * us) + rx1 * ex) + r2ypz * fa) + r2ymz * fb) + c.mc;
        cpx<R> ey = ((((((lp2u(1) - lp2u(2)) + lp2u(5)) - lp2u(6)) + lp2u(7)) - lp2u(8)) + lp2u(3)) - lp2u(4);
        fa = ((c.U(1) - c.U(8)) + c.U(2)) - c.U(7);
        fb = ((c.U(6) - c.U(3)) + c.U(5)) - c.U(4);
        t1 = ((((-(ry2 * ls) - rxpz * us) + ry1 * ey) + r2xpz * fa) + r2xmz * fb) + c.mc;
        cpx<R> ez = ((((((lp2u(1) - lp2u(8)) + lp2u(2)) - lp2u(7)) + lp2u(4)) - lp2u(5)) + lp2u(3)) - lp2u(6);
        fa = ((c.U(1) - c.U(8)) + c.U(5)) - c.U(4);
        fb = ((c.U(2) - c.U(7)) + c.U(6)) - c.U(3);
        t2 = ((((-(rz2 * ls) - rxpy * us) + rz1 * ez) + r2xpy * fa) + r2xmy * fb) + c.mc;
        t3 = rxy * (((((((lpu(2) - lpu(1)) + lpu(7)) - lpu(8)) + lpu(3)) - lpu(4)) + lpu(6)) - lpu(5));
        t4 = rxz * (((((((lpu(4) - lpu(1)) + lpu(5)) - lpu(8)) + lpu(3)) - lpu(2)) + lpu(6)) - lpu(7));
        t5 = ryz * (((((((lpu(2) - lpu(1)) + lpu(7)) - lpu(8)) + lpu(4)) - lpu(3)) + lpu(5)) - lpu(6));
        put(t0, t1, t2, t3, t3, t4, t4, t5, t5);
        break;
    }
    case 14:  // a6n
        t0 = rz2 * (((c.U(8) + c.U(5)) + c.U(7)) + c.U(6)) + m2;
        t1 = rz2 * (((lp2u(8) + lp2u(5)) + lp2u(7)) + lp2u(6)) + m2;
        t2 = -(rxy * (((lpu(8) + lpu(5)) - lpu(7)) - lpu(6)));
        t3 = rxz * (((lmu(8) - lmu(5)) + lmu(7)) - lmu(6));
        t4 = ryz * (((lmu(8) - lmu(5)) + lmu(6)) - lmu(7));
        put(t0, t0, t1, t2, t2, t3, -t3, t4, -t4);
        break;
    case 15:  // a7n
        t0 = -(ry2 * (c.U(4) + c.U(3))) + m3;
        t2 = -(rxy * (lmu(4) - lmu(3))); t3 = -(rxz * (lpu(4) + lpu(3))); t4 = -(ryz * (lmu(4) - lmu(3)));
        put(t0, t0, t0, t2, -t2, t3, t3, t4, -t4);
        break;
    case 16:  // a8n
        t0 = rx2 * (((lp2u(8) + lp2u(3)) + lp2u(7)) + lp2u(4)) + m2;
        t1 = rx2 * (((c.U(8) + c.U(3)) + c.U(7)) + c.U(4)) + m2;
        t2 = rxy * (((lmu(7) - lmu(4)) + lmu(3)) - lmu(8));
        t3 = -(rxz * (((lmu(7) - lmu(4)) + lmu(8)) - lmu(3)));
        t4 = ryz * (((lpu(7) + lpu(4)) - lpu(8)) - lpu(3));
        put(t0, t0, t1, t2, -t2, t3, -t3, t4, t4);
        break;
    case 17:  // a9n
        t0 = -(ry2 * (c.U(8) + c.U(7))) + m3;
        t2 = rxy * (lmu(7) - lmu(8)); t3 = rxz * (lpu(8) + lpu(7)); t4 = ryz * (lmu(8) - lmu(7));
        put(t0, t0, t0, t2, -t2, t3, t3, t4, -t4);
        break;
    // ---------------- oy = +1 ----------------
    case 18:  // a1p
        t0 = -(rxy * lpu(2)); t1 = rxz * lpu(2); t2 = -(ryz * lpu(2));
        put(corner(rx2, rypz, -1, rx1, +1, r2ymz, 2, m4), corner(ry2, rxpz, +1, ry1, -1, r2xpz, 2, m4),
            corner(rz2, rxpy, -1, rz1, -1, r2xmy, 2, m4), t0, t0, t1, t1, t2, t2);
        break;
    case 19:  // a2p
        t0 = -(rz2 * (c.U(2) + c.U(6))) + m3; t1 = -(rz2 * (lp2u(2) + lp2u(6))) + m3;
        t2 = -(rxy * (lpu(2) + lpu(6))); t3 = -(rxz * (lmu(2) - lmu(6))); t4 = -(ryz * (lmu(6) - lmu(2)));
        put(t0, t0, t1, t2, t2, t3, -t3, t4, -t4);
        break;
    case 20:  // a3p
        t0 = -(rxy * lpu(6)); t1 = -(rxz * lpu(6)); t2 = ryz * lpu(6);
        put(corner(rx2, rypz, -1, rx1, +1, r2ypz, 6, m4), corner(ry2, rxpz, +1, ry1, -1, r2xmz, 6, m4),
            corner(rz2, rxpy, +1, rz1, -1, r2xpy, 6, m4), t0, t0, t1, t1, t2, t2);
        break;
    case 21:  // a4p
        t0 = -(rx2 * (lp2u(2) + lp2u(4))) + m3; t1 = -(rx2 * (c.U(2) + c.U(4))) + m3;
        t2 = rxy * (lmu(4) - lmu(2)); t3 = rxz * (lmu(2) - lmu(4)); t4 = -(ryz * (lpu(2) + lpu(4)));
        put(t0, t0, t1, t2, -t2, t3, -t3, t4, t4);
        break;
    case 22:  // a5p
        t0 = ry2 * (((c.U(8) + c.U(2)) + c.U(4)) + c.U(6)) + m2;
        t2 = rxy * (((lmu(8) - lmu(2)) + lmu(4)) - lmu(6));
        t3 = -(rxz * (((lpu(8) + lpu(2)) - lpu(4)) - lpu(6)));
        t4 = -(ryz * (((lmu(8) - lmu(2)) - lmu(4)) + lmu(6)));
        put(t0, t0, t0, t2, -t2, t3, t3, t4, -t4);
        break;
    case 23:  // a6p
        t0 = -(rx2 * (lp2u(8) + lp2u(6))) + m3; t1 = -(rx2 * (c.U(8) + c.U(6))) + m3;
        t2 = rxy * (lmu(8) - lmu(6)); t3 = rxz * (lmu(8) - lmu(6)); t4 = ryz * (lpu(8) + lpu(6));
        put(t0, t0, t1, t2, -t2, t3, -t3, t4, t4);
        break;
    case 24:  // a7p
        t0 = rxy * lpu(4); t1 = -(rxz * lpu(4)); t2 = -(ryz * lpu(4));
        put(corner(rx2, rypz, +1, rx1, +1, r2ymz, 4, m4), corner(ry2, rxpz, +1, ry1, +1, r2xmz, 4, m4),
            corner(rz2, rxpy, -1, rz1, +1, r2xmy, 4, m4), t0, t0, t1, t1, t2, t2);
        break;
    case 25:  // a8p
        t0 = -(rz2 * (c.U(8) + c.U(4))) + m3; t1 = -(rz2 * (lp2u(8) + lp2u(4))) + m3;
        t2 = rxy * (lpu(8) + lpu(4)); t3 = -(rxz * (lmu(8) - lmu(4))); t4 = -(ryz * (lmu(8) - lmu(4)));
        put(t0, t0, t1, t2, t2, t3, -t3, t4, -t4);
        break;
    default:  // 26, a9p
        t0 = rxy * lpu(8); t1 = rxz * lpu(8); t2 = ryz * lpu(8);
        put(corner(rx2, rypz, +1, rx1, +1, r2ypz, 8, m4), corner(ry2, rxpz, +1, ry1, +1, r2xpz, 8, m4),
            corner(rz2, rxpy, +1, rz1, +1, r2xpy, 8, m4), t0, t0, t1, t1, t2, t2);
        break;
    }
}

// The eight corner labels (all three offsets nonzero) share one formula that
// differs only in the octant cell and in signs, so one pass can evaluate a
// different corner on every 16 lanes.  kCornerK[j] is the label of corner j;
// kCorner[j] packs its cell (bits 0-3) and, set for +, the signs of the
// xy, xz, yz off-diagonal terms (bits 4-6) and, per diagonal entry x, y, z,
// of the ra and rb terms and of the second r1 term inside rb (bits 7-15).
// The table was read off label_block's corner cases; corner_block computes
// exactly what those cases compute, operation for operation.
__constant__ const int kCornerK[8] = {0, 2, 6, 8, 18, 20, 24, 26};
__constant__ const uint32_t kCorner[8] = {0x9271, 0x2015, 0xcac3, 0x78a7, 0x1522, 0xa746, 0x4d94, 0xfff8};

template <typename R, typename F, typename E>
__device__ __forceinline__ void corner_block(uint32_t prm, const cpx<R>* cup, const cpx<R>* cdn, F get, R wm4o8,
                                             R ro, E emit)
{
    // get(i): the node's r-coefficient i (0..8 = rx2, ry2, rz2, rxy, rxz, ryz,
    // rx1, ry1, rz1), fetched where it is used; emit(e, value) takes entry e
    // of the row-major 3x3 block as soon as it is formed
    const int cell = (int)(prm & 15);
    auto pos = [&](int b) { return ((prm >> (4 + b)) & 1) != 0; };
    const cpx<R> U = cell <= 4 ? cup[cell - 1] : cdn[cell - 5];
    const cpx<R> D = cell <= 4 ? cup[3 + cell] : cdn[cell - 1];
    const cpx<R> L = (D + U) + U, P = D + U;
    const cpx<R> m4(wm4o8 * ro, R(0));
    auto corner = [&](cpx<R> q1, cpx<R> qs, bool sa, cpx<R> ra, bool sb, cpx<R> rb) {
        cpx<R> s = q1 * L + qs * U;
        cpx<R> ta = ra * L, tb = rb * U;
        s = s + (sa ? ta : -ta);
        s = s + (sb ? tb : -tb);
        return s + m4;
    };
    {
        const cpx<R> rx2 = get(0), ry2 = get(1), rz2 = get(2), ry1 = get(7), rz1 = get(8);
        const cpx<R> rbx = pos(5) ? ry1 + rz1 : ry1 - rz1;
        emit(0, corner(rx2, ry2 + rz2, pos(3), get(6), pos(4), rbx));
    }
    {
        const cpx<R> rx2 = get(0), ry2 = get(1), rz2 = get(2), rx1 = get(6), rz1 = get(8);
        const cpx<R> rby = pos(8) ? rx1 + rz1 : rx1 - rz1;
        emit(4, corner(ry2, rx2 + rz2, pos(6), get(7), pos(7), rby));
    }
    {
        const cpx<R> rx2 = get(0), ry2 = get(1), rz2 = get(2), rx1 = get(6), ry1 = get(7);
        const cpx<R> rbz = pos(11) ? rx1 + ry1 : rx1 - ry1;
        emit(8, corner(rz2, rx2 + ry2, pos(9), get(8), pos(10), rbz));
    }
    cpx<R> t0 = get(3) * P;
    if (!pos(0)) t0 = -t0;
    emit(1, t0);
    emit(3, t0);
    cpx<R> t1 = get(4) * P;
    if (!pos(1)) t1 = -t1;
    emit(2, t1);
    emit(6, t1);
    cpx<R> t2 = get(5) * P;
    if (!pos(2)) t2 = -t2;
    emit(5, t2);
    emit(7, t2);
}

// The twelve edge labels (two nonzero offsets) fall into three families by
// their zero offset -- z (labels 1, 7, 19, 25), x (3, 5, 21, 23), y (9, 11,
// 15, 17) -- each with one formula that differs only in the two octant cells,
// the signs of the xy, xz, yz terms and the order of the differences in them.
// kEdge[j] packs cells a, b (bits 0-3, 4-7), the signs, set for + (bits
// 8-10), and, set when a difference is lmu(a) - lmu(b), its order (bits
// 11-13).  Read off label_block's edge cases; edge_block computes exactly
// what those cases compute.
__constant__ const int kEdgeK[12] = {1, 7, 19, 25, 3, 5, 21, 23, 9, 11, 15, 17};
__constant__ const uint32_t kEdge[12] = {0x3951, 0x1837, 0x1862, 0x3948, 0x3f31, 0x3357,
                                         0x3342, 0x3f68, 0x3721, 0x3865, 0x3834, 0x3778};

// F: the family, 0 = z, 1 = x, 2 = y; get and emit as for corner_block.
template <typename R, int F, typename G, typename E>
__device__ __forceinline__ void edge_block(uint32_t prm, const cpx<R>* cup, const cpx<R>* cdn, G get, R wm3o12,
                                           R ro, E emit)
{
    const int a = (int)(prm & 15), b = (int)((prm >> 4) & 15);
    auto pos = [&](int bit) { return ((prm >> bit) & 1) != 0; };
    const cpx<R> Ua = a <= 4 ? cup[a - 1] : cdn[a - 5], Ub = b <= 4 ? cup[b - 1] : cdn[b - 5];
    const cpx<R> Da = a <= 4 ? cup[3 + a] : cdn[a - 1], Db = b <= 4 ? cup[3 + b] : cdn[b - 1];
    const cpx<R> m3(wm3o12 * ro, R(0));
    const cpx<R> lmuA = Da - Ua, lmuB = Db - Ub;
    auto diff = [&](int ordbit) { return pos(ordbit) ? lmuA - lmuB : lmuB - lmuA; };
    auto sgn = [&](int bit, cpx<R> x) { return pos(bit) ? x : -x; };
    {
        const cpx<R> rF2 = get(F == 0 ? 2 : (F == 1 ? 0 : 1));  // rz2, rx2, ry2
        const cpx<R> tU = -(rF2 * (Ua + Ub)) + m3;
        if (F == 2) {
            emit(0, tU); emit(4, tU); emit(8, tU);
        } else {
            const cpx<R> tL = -(rF2 * (((Da + Ua) + Ua) + ((Db + Ub) + Ub))) + m3;
            if (F == 0) { emit(0, tU); emit(4, tU); emit(8, tL); }
            else { emit(0, tL); emit(4, tL); emit(8, tU); }
        }
    }
    const cpx<R> S = (Da + Ua) + (Db + Ub);
    const cpx<R> t2 = sgn(8, get(3) * (F == 0 ? S : diff(11)));
    emit(1, t2); emit(3, F == 0 ? t2 : -t2);
    const cpx<R> t3 = sgn(9, get(4) * (F == 2 ? S : diff(12)));
    emit(2, t3); emit(6, F == 2 ? t3 : -t3);
    const cpx<R> t4 = sgn(10, get(5) * (F == 1 ? S : diff(13)));
    emit(5, t4); emit(7, F == 1 ? t4 : -t4);
}

// The six face labels (one nonzero offset) pair up by that offset's axis --
// y (labels 4, 22), x (10, 16), z (12, 14) -- with one formula per pair that
// differs only in the octant cells and their order in each four-term sum.
// kFace[j] packs three lists of four cells, 4 bits each (list l, entry i at
// bits 16 l + 4 i): y uses list 0 throughout; x list 0 for the diagonal, 1 for
// the xy term, 2 for xz and yz; z list 0, and list 2 for the yz term.  Read
// off label_block's face cases; face_block computes exactly what those cases
// compute.
__constant__ const int kFaceK[6] = {4, 22, 10, 16, 12, 14};
__constant__ const uint64_t kFace[6] = {0x357135713571ull, 0x642864286428ull, 0x615216525261ull,
                                        0x384783474738ull, 0x234132413241ull, 0x765867586758ull};

// F: the axis, 0 = y, 1 = x, 2 = z; get and emit as for corner_block.
template <typename R, int F, typename G, typename E>
__device__ __forceinline__ void face_block(uint64_t cells, const cpx<R>* cup, const cpx<R>* cdn, G get, R wm2o6,
                                           R ro, E emit)
{
    auto cell = [&](int l, int i) { return (int)((cells >> (16 * l + 4 * i)) & 15); };
    auto Uc = [&](int x) { return x <= 4 ? cup[x - 1] : cdn[x - 5]; };
    auto Dc = [&](int x) { return x <= 4 ? cup[3 + x] : cdn[x - 1]; };
    auto U = [&](int l, int i) { return Uc(cell(l, i)); };
    auto lpu = [&](int l, int i) { const int x = cell(l, i); return Dc(x) + Uc(x); };
    auto lmu = [&](int l, int i) { const int x = cell(l, i); return Dc(x) - Uc(x); };
    auto lp2u = [&](int l, int i) { const int x = cell(l, i); return (Dc(x) + Uc(x)) + Uc(x); };
    const cpx<R> m2(wm2o6 * ro, R(0));
    if (F == 0) {
        const cpx<R> t0 = get(1) * (((U(0, 0) + U(0, 1)) + U(0, 2)) + U(0, 3)) + m2;
        emit(0, t0); emit(4, t0); emit(8, t0);
        const cpx<R> t2 = get(3) * (((lmu(0, 0) - lmu(0, 1)) + lmu(0, 2)) - lmu(0, 3));
        emit(1, t2); emit(3, -t2);
        const cpx<R> t3 = -(get(4) * (((lpu(0, 0) + lpu(0, 1)) - lpu(0, 2)) - lpu(0, 3)));
        emit(2, t3); emit(6, t3);
        const cpx<R> t4 = -(get(5) * (((lmu(0, 0) - lmu(0, 1)) - lmu(0, 2)) + lmu(0, 3)));
        emit(5, t4); emit(7, -t4);
    } else if (F == 1) {
        const cpx<R> rx2 = get(0);
        const cpx<R> t0 = rx2 * (((lp2u(0, 0) + lp2u(0, 1)) + lp2u(0, 2)) + lp2u(0, 3)) + m2;
        const cpx<R> t1 = rx2 * (((U(0, 0) + U(0, 1)) + U(0, 2)) + U(0, 3)) + m2;
        emit(0, t0); emit(4, t0); emit(8, t1);
        const cpx<R> t2 = get(3) * (((lmu(1, 0) - lmu(1, 1)) + lmu(1, 2)) - lmu(1, 3));
        emit(1, t2); emit(3, -t2);
        const cpx<R> t3 = -(get(4) * (((lmu(2, 0) - lmu(2, 1)) + lmu(2, 2)) - lmu(2, 3)));
        emit(2, t3); emit(6, -t3);
        const cpx<R> t4 = get(5) * (((lpu(2, 0) + lpu(2, 1)) - lpu(2, 2)) - lpu(2, 3));
        emit(5, t4); emit(7, t4);
    } else {
        const cpx<R> rz2 = get(2);
        const cpx<R> t0 = rz2 * (((U(0, 0) + U(0, 1)) + U(0, 2)) + U(0, 3)) + m2;
        const cpx<R> t1 = rz2 * (((lp2u(0, 0) + lp2u(0, 1)) + lp2u(0, 2)) + lp2u(0, 3)) + m2;
        emit(0, t0); emit(4, t0); emit(8, t1);
        const cpx<R> t2 = -(get(3) * (((lpu(0, 0) + lpu(0, 1)) - lpu(0, 2)) - lpu(0, 3)));
        emit(1, t2); emit(3, t2);
        const cpx<R> t3 = get(4) * (((lmu(0, 0) - lmu(0, 1)) + lmu(0, 2)) - lmu(0, 3));
        emit(2, t3); emit(6, -t3);
        const cpx<R> t4 = get(5) * (((lmu(2, 0) - lmu(2, 1)) + lmu(2, 2)) - lmu(2, 3));
        emit(5, t4); emit(7, -t4);
    }
}

// Values kernel geometry: a block of ASM_THREADS threads owns a segment of
// ASM_WAVE z-nodes of an (x, y) column and writes it in chunks of CH nodes.
constexpr int CH = 16;
constexpr int ASM_THREADS = 256;
constexpr int TZ = ASM_WAVE + 2;  // tile z-slots of a segment (one node either side)

// Entries before 0-based z-node m of an (x, y) column's rows, in units of
// 9 cxy: a node has 3 rows of 3 cxy czn entries, czn = 2 at both ends of the
// column and 3 inside.
__device__ inline int col_units(int m, int nz) { return 3 * m - (m >= 1) - (m >= nz); }

// Block id -> (z segment, x, y), with consecutive ids of one XCD (blocks go
// round-robin to the 8 XCDs) mapped to consecutive columns, so the blocks
// sharing an L2 share input columns and write adjacent output.
__device__ inline void segment_of_block(int64_t nblk, int nzs, int nx, int& zs, int& ix, int& iy)
{
    int64_t b = blockIdx.x;
    int64_t q = nblk / 8, r = nblk % 8, x = b % 8, i = b / 8;
    b = x * q + (x < r ? x : r) + i;
    zs = (int)(b % nzs);
    int64_t c = b / nzs;
    ix = (int)(c % nx) + 1;
    iy = (int)(c / nx) + 1;
}

// Region class (fdfd3d.f:50-115) and per-node coefficients of node (ix, iy,
// iz) whose tile z-slot is tz: mass weights (calcwts.f:28-71, with the
// moduli of setcoefs3d.f:224-225) and PML stretch factors (pml3d.f:207-295);
// the octant cells come from the segment's cell table s_cell.  Returns 0
// interior, 1 PML, 2 Dirichlet, 3 traction-free top (tpfree.f, iz = 1 with
// usemin = false; then rx1, ry1, rz1 hold r1x dxo8, r1y dxo8, r1z dxo8).
template <typename R, int TZN, int CW>
__device__ inline int node_setup(const wt_grid_params& p, int ix, int iy, int iz, int tz,
                                 const cpx<R> (*s_lam)[TZN], const cpx<R> (*s_mu)[TZN],
                                 const R (*s_rho)[TZN], const cpx<R> (*s_cell)[CW], int* err, bool report,
                                 NodeCoef<R>& c, bool& topfree)
{
    const int nx = p.nx, ny = p.ny, nz = p.nz, ip = p.ipml;
    const bool inx = ix > ip && ix <= nx - ip, iny = iy > ip && iy <= ny - ip;
    int mode = 1, izp = iz;
    if (inx && iny && iz > ip && iz <= nz - ip) mode = 0;
    else if (iz == nz || ix == 1 || ix == nx || iy == 1 || iy == ny) mode = 2;
    else if (iz <= ip) {
        if (iz == 1 && !p.fst) mode = 2;
        else if (iz == 1 && !p.usemin) mode = 3;
        else if (!p.fst) mode = 1;
        else if (inx && iny) mode = 0;
        else { mode = 1; izp = ip + 1; }
    }
    topfree = p.fst && iz == 1;

    c.cup = s_cell[tz - 1];
    c.cdn = s_cell[tz];
    cpx<R> lc = s_lam[4][tz], mcn = s_mu[4][tz];
    R rc = s_rho[4][tz];
    R vp = csqrt_re((lc + R(2) * mcn) / rc);
    R vs = csqrt_re(mcn / rc);
    double x = (double)(vp / vs);
    double w1 = 0.2909366e+00 + x * (-0.1792402e-01 + x * (0.7329184e-02 + x * -0.1055034e-02));
    double w2 = 0.0, w3 = 0.0;
    if (w1 > 1.0) w1 = 1.0;
    else w2 = 0.6367377e+00 + x * (-0.7982022e-02 + x * (0.3385164e-02 + x * -0.5065399e-03));
    if ((w1 + w2) > 1.0) w2 = 0.0;
    else w3 = 0.7244875e-01 + x * (0.2569959e-01 + x * (-0.1060011e-01 + x * 0.1540534e-02));
    if ((w1 + w2 + w3) > 1.0) w3 = 1.0 - (w1 + w2);
    double w4 = 1.0 - (w1 + w2 + w3);
    const R wm1 = (R)w1, wm2 = (R)w2, wm3 = (R)w3, wm4 = (R)w4;

    const cpx<R> om(p.omega_re, p.omega_im);
    const cpx<R> eyeom(om.im, -om.re);
    const cpx<R> osq = om * om;
    c.wm2o6 = (osq.re * wm2) / R(6);
    c.wm3o12 = (osq.re * wm3) / R(12);
    c.wm4o8 = (osq.re * wm4) / R(8);
    cpx<R> mcen = wm1 * osq;
    c.mc = cpx<R>(mcen.re * rc, mcen.im * rc);

    const R dx = (R)p.dx, dx2 = dx + dx, dxx = dx * dx, dxx4 = (dxx + dxx) + (dxx + dxx);
    cpx<R> r1x(1, 0), r1y(1, 0), r1z(1, 0), r2x(0, 0), r2y(0, 0), r2z(0, 0);
    if (mode == 1) {
        const R pmld = (R)p.pmld, pmlf = (R)p.pmlf;
        bool ok = pml_axis(ix, nx, ip, pmld, dx, pmlf, vp, eyeom, r1x, r2x);
        ok = pml_axis(iy, ny, ip, pmld, dx, pmlf, vp, eyeom, r1y, r2y) && ok;
        ok = pml_axis(izp, nz, ip, pmld, dx, pmlf, vp, eyeom, r1z, r2z) && ok;
        if (!ok && report) atomicOr(err, 1);
    }
    if (mode == 3) {
        // tpfree.f:67-124: the x and y stretch factors r1 of the PML form
        // (kmax = 1, fdom = 0), r1z = 1, scaled by 1 / (8 dx)
        const R pmld = (R)p.pmld, pmlf = (R)p.pmlf;
        bool ok = pml_axis(ix, nx, ip, pmld, dx, pmlf, vp, eyeom, r1x, r2x);
        ok = pml_axis(iy, ny, ip, pmld, dx, pmlf, vp, eyeom, r1y, r2y) && ok;
        if (!ok && report) atomicOr(err, 1);
        const R dxo8 = R(1) / (dx * R(8));
        c.rx1 = dxo8 * r1x;
        c.ry1 = dxo8 * r1y;
        c.rz1 = dxo8 * r1z;
    } else {
        c.rx1 = (r2x / dx2) / R(4);
        c.ry1 = (r2y / dx2) / R(4);
        c.rz1 = (r2z / dx2) / R(4);
    }
    c.rx2 = (((r1x * r1x) / dxx) / R(4)) / R(2);
    c.ry2 = (((r1y * r1y) / dxx) / R(4)) / R(2);
    c.rz2 = (((r1z * r1z) / dxx) / R(4)) / R(2);
    c.rxy = ((r1x * r1y) / dxx4) / R(4);
    c.rxz = ((r1x * r1z) / dxx4) / R(4);
    c.ryz = ((r1y * r1z) / dxx4) / R(4);
    return mode;
}

// The traction-free top (tpfree.f:55-420, fdfd3d.f:73-85): at iz = 1 with
// usemin = false the node couples only to its four z+1 corners, through the
// octant cell below the node on that side (5..8), and to itself; no mass
// term.  c.rx1, c.ry1, c.rz1 hold r1x/(8 dx), r1y/(8 dx), r1z/(8 dx).  The
// centre's VW entry takes the lambda cells where UW takes mu, as
// tpfree.f:321 has it.  Writes the five 3x3 blocks of node (1, ix, iy) into
// the chunk span that starts at that node (a non-Dirichlet node: every x, y
// neighbour exists).
template <typename R, typename V>
__device__ void tpfree_store(const NodeCoef<R>& c, int ix, int iy, int nx, int ny, int nz, V* stage)
{
    const cpx<R> X = c.rx1, Y = c.ry1, Z = c.rz1;
    auto lp2u = [&](int i) { return (c.D(i) + c.U(i)) + c.U(i); };
    const int cxn = 1 + (ix > 1) + (ix < nx), cyn = 1 + (iy > 1) + (iy < ny);
    const int czn = 1 + (nz > 1);  // iz = 1: itself and z + 1
    const int rl = 3 * cxn * cyn * czn;
    auto put = [&](int ox, int oy, int oz, const cpx<R> b[9]) {
        const int ry = oy > -1 ? (iy > 1) + (oy > 0) : 0, rx = ox > -1 ? (ix > 1) + (ox > 0) : 0;
        const int pos = ((ry * cxn + rx) * czn + oz) * 3;
        for (int r = 0; r < 3; ++r)
            for (int e = 0; e < 3; ++e) {
                V w;
                w.x = b[3 * r + e].re;
                w.y = b[3 * r + e].im;
                stage[pos + r * rl + e] = w;
            }
    };
    const cpx<R> z0(R(0), R(0));
    // corners (ox, oy) with cells 5 (-,-), 6 (-,+), 7 (+,-), 8 (+,+)
    for (int q = 0; q < 4; ++q) {
        const int ox = q < 2 ? -1 : 1, oy = (q & 1) ? 1 : -1, cell = 5 + q;
        const cpx<R> u = c.U(cell), d = c.D(cell);
        const cpx<R> xu = X * u, yu = Y * u, xd = X * d, yd = Y * d;
        const cpx<R> b[9] = {Z * u, z0, ox > 0 ? xu : -xu,
                             z0, Z * u, oy > 0 ? yu : -yu,
                             ox > 0 ? xd : -xd, oy > 0 ? yd : -yd, Z * lp2u(cell)};
        put(ox, oy, 1, b);
    }
    const cpx<R> su = ((c.U(8) + c.U(7)) + c.U(6)) + c.U(5);
    const cpx<R> sl = ((lp2u(8) + lp2u(7)) + lp2u(6)) + lp2u(5);
    const cpx<R> dy = ((c.D(5) - c.D(8)) - c.D(6)) + c.D(7);
    const cpx<R> b[9] = {-(Z * su), z0, X * (((c.U(5) - c.U(8)) + c.U(6)) - c.U(7)),
                         z0, -(Z * su), Y * dy,
                         X * (((c.D(5) - c.D(8)) + c.D(6)) - c.D(7)), Y * dy, -(Z * sl)};
    put(0, 0, 0, b);
}

// Marks every coefficient as possibly changed here (an empty asm statement).
template <typename R>
__device__ __forceinline__ void launder(NodeCoef<R>& c)
{
    R* f = reinterpret_cast<R*>(&c);
#pragma unroll
    for (int i = 0; i < (int)(__builtin_offsetof(NodeCoef<R>, cup) / sizeof(R)); ++i) asm volatile("" : "+v"(f[i]));
}

// Neighbour labels of each wave of a block, 6 bits per turn: a label 0..26
// runs alone (the label uniform over the wave); kQuad + j runs a group at
// once, one label per 16 lanes -- j = 0, 1: corners 0-3, 4-7 of kCornerK;
// j = 2, 3, 4: the z, x, y edge families of kEdgeK; j = 5, 6, 7: the y, x, z
// face pairs of kFace (32 lanes) -- and kEnd ends the list.  The groups and
// the centre are spread so the four waves reach the chunk's barrier about
// together (measured per wave).
constexpr int kQuad = 32, kEnd = 63;
constexpr uint64_t lbl7(int a, int b, int c, int d, int e, int f, int g)
{
    return (uint64_t)a | (uint64_t)b << 6 | (uint64_t)c << 12 | (uint64_t)d << 18 | (uint64_t)e << 24 |
           (uint64_t)f << 30 | (uint64_t)g << 36;
}
__constant__ const uint64_t kWaveLabels[4] = {
    lbl7(kQuad + 6, kQuad + 4, kEnd, kEnd, kEnd, kEnd, kEnd), lbl7(13, kQuad + 0, kEnd, kEnd, kEnd, kEnd, kEnd),
    lbl7(kQuad + 5, kQuad + 1, kEnd, kEnd, kEnd, kEnd, kEnd), lbl7(kQuad + 2, kQuad + 3, kQuad + 7, kEnd, kEnd, kEnd, kEnd)};

// The rows of a chunk of CH z-nodes are one contiguous span of the CSR values
// (node after node, 3 rows each, neighbours ascending in y, x, z), so the
// block builds the span in LDS and writes it front to back with consecutive
// threads on consecutive entries: every store instruction covers whole cache
// lines and the block's output is one sequential sweep.
//
// A block owns a segment of ASM_WAVE z-nodes of one (x, y) column; lane l of
// every wave holds node l of the segment.  Wave 0 forms the node
// coefficients once for all chunks (their divisions and square roots are not
// repeated) and the other waves take them from LDS.  The
// segment is then written chunk by chunk: each wave evaluates its labels of
// kWaveLabels for the chunk's 16 nodes, and the span goes out while the next
// chunk is evaluated.
template <typename R>
__global__ __launch_bounds__(ASM_THREADS, 2) void assemble_values_kernel(
    wt_grid_params p, const cpx<R>* __restrict__ lam, const cpx<R>* __restrict__ mu,
    const R* __restrict__ rho, cpx<R>* __restrict__ out, int* __restrict__ err, int64_t nblk)
{
    // element type moved as one vector: 16 bytes (FP64), 8 bytes (FP32)
    using V = typename std::conditional<sizeof(R) == 8, double2, float2>::type;
    constexpr int STG_N = CH * 3 * 81;
    constexpr int IN_BYTES = 2 * 9 * TZ * (int)sizeof(cpx<R>);
    constexpr int STG_BYTES = STG_N * (int)sizeof(V);
    // The lambda/mu tiles are dead once the node coefficients are formed; the
    // span reuses their space.
    __shared__ __align__(16) char s_raw[IN_BYTES > STG_BYTES ? IN_BYTES : STG_BYTES];
    __shared__ R s_rho[9][TZ];
    // Octant cells between tile z-slots j and j + 1: mu of the four (x, y)
    // quadrants, then lambda; one pad entry keeps the lanes' rows apart in
    // the LDS banks.
    constexpr int CW = 9;
    __shared__ __align__(16) cpx<R> s_cell[TZ - 1][CW];
    cpx<R>(*s_lam)[TZ] = (cpx<R>(*)[TZ])s_raw;
    cpx<R>(*s_mu)[TZ] = (cpx<R>(*)[TZ])(s_raw + 9 * TZ * sizeof(cpx<R>));
    V* const stage = (V*)s_raw;
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    int zs, ix, iy;
    segment_of_block(nblk, (nz + ASM_WAVE - 1) / ASM_WAVE, nx, zs, ix, iy);
    const int t = threadIdx.x;
    const int n = t & (ASM_WAVE - 1);                        // node of the segment
    const int wv = __builtin_amdgcn_readfirstlane(t >> 6);  // wave of the block
    const int z0 = zs * ASM_WAVE;                            // 0-based first z of the segment
    const int zend = nz < z0 + ASM_WAVE ? nz : z0 + ASM_WAVE;

    // Tile slot zz holds 1-based z = z0 + zz, clamped to the grid, of the
    // clamped 3x3 column neighbourhood (column index 3 (x offset) + y offset).
    for (int i = t; i < 9 * TZ; i += ASM_THREADS) {
        const int cc = i / TZ, zz = i - cc * TZ;
        int xg = ix + cc / 3 - 1, yg = iy + cc % 3 - 1, zg = z0 + zz;
        xg = xg < 1 ? 1 : (xg > nx ? nx : xg);
        yg = yg < 1 ? 1 : (yg > ny ? ny : yg);
        zg = zg < 1 ? 1 : (zg > nz ? nz : zg);
        const int64_t gi = (int64_t)(xg - 1) * nz + (int64_t)(yg - 1) * nx * nz + zg - 1;
        s_lam[cc][zz] = lam[gi];
        s_mu[cc][zz] = mu[gi];
        s_rho[cc][zz] = rho[gi];
    }
    __syncthreads();
    const int m = z0 + n;  // 0-based z of this lane's node
    const bool live = m < zend;
    const int iz = live ? m + 1 : zend, tz = iz - z0;
    // Wave 0 forms the node coefficients and leaves them in LDS past the
    // lambda/mu tiles (space the span takes over later) while the other three
    // waves build the octant-cell table; the other waves then read the
    // coefficients instead of forming them again.
    constexpr int NF = (int)(__builtin_offsetof(NodeCoef<R>, cup) / sizeof(R));
    static_assert(IN_BYTES + (NF + 1) * ASM_WAVE * (int)sizeof(R) <= STG_BYTES, "coefficients overflow the span");
    R* const s_coef = (R*)(s_raw + IN_BYTES);  // [NF][ASM_WAVE], then the mode word per node
    int* const s_mode = (int*)(s_coef + NF * ASM_WAVE);
    NodeCoef<R> c;
    bool topfree;
    int mode;
    if (wv == 0) {
        mode = node_setup<R, TZ, CW>(p, ix, iy, iz, tz, s_lam, s_mu, s_rho, s_cell, err, live, c, topfree);
        const R* f = reinterpret_cast<const R*>(&c);
#pragma unroll
        for (int i = 0; i < NF; ++i) s_coef[i * ASM_WAVE + n] = f[i];
        s_mode[n] = mode | (topfree ? 4 : 0);
    } else {
        // Octant-cell moduli (setcoefs3d.f:108-174): the mean of a quadrant's
        // four columns at two z-levels.  A node's cells 1..4 are the interface
        // above it and 5..8 the one below, so each interface is formed once
        // (the two nodes sharing it add the same two terms, in either order).
        // At the free surface the cells above the top node are zero.
        for (int i = t - ASM_WAVE; i < (TZ - 1) * 4; i += ASM_THREADS - ASM_WAVE) {
            const int j = i >> 2, q = i & 3;
            const int c0 = (q & 1) ? ((q & 2) ? 8 : 2) : ((q & 2) ? 6 : 0);
            const int c1 = (q & 1) ? 5 : 3, c2 = (q & 2) ? 7 : 1;
            auto km = [&](const cpx<R>(*a)[TZ], int z) {
                return (((a[c0][z] + a[c1][z]) + a[c2][z]) + a[4][z]) / R(4);
            };
            cpx<R> uu = (km(s_mu, j) + km(s_mu, j + 1)) / R(2);
            cpx<R> dd = (km(s_lam, j) + km(s_lam, j + 1)) / R(2);
            if (j == 0 && z0 == 0 && p.fst) { uu = cpx<R>(0, 0); dd = cpx<R>(0, 0); }
            s_cell[j][q] = uu;
            s_cell[j][4 + q] = dd;
        }
    }
    __syncthreads();
    if (wv != 0) {
        R* f = reinterpret_cast<R*>(&c);
#pragma unroll
        for (int i = 0; i < NF; ++i) f[i] = s_coef[i * ASM_WAVE + n];
        const int mt = s_mode[n];
        mode = mt & 3;
        topfree = (mt & 4) != 0;
        c.cup = s_cell[tz - 1];
        c.cdn = s_cell[tz];
    }
    const int cxn = 1 + (ix > 1) + (ix < nx), cyn = 1 + (iy > 1) + (iy < ny), cxy = cxn * cyn;
    const int czn = 1 + (iz > 1) + (iz < nz);
    const int rl = 3 * cxy * czn;
    V* const col = (V*)(out + node_offset(1, ix, iy, nx, ny, nz));
    __syncthreads();  // lambda/mu tiles and coefficient copies dead from here on

#pragma unroll 1
    for (int m0 = z0; m0 < zend; m0 += CH) {
        const int m1 = m0 + CH < zend ? m0 + CH : zend;  // chunk = 0-based z-nodes [m0, m1)
        const bool mine = live && m >= m0 && m < m1;
        const int base = 9 * cxy * (col_units(m, nz) - col_units(m0, nz));
#pragma unroll 1
        for (int turn = 0; turn < 7; ++turn) {
            const int k = (int)((kWaveLabels[wv] >> (6 * turn)) & 63);
            if (k == kEnd) break;
            // Keep the compiler from hoisting any label's subexpressions out
            // of the loop (all live at once, they spill): the node
            // coefficients count as redefined on each pass.
            launder(c);
            if (k >= kQuad) {
                // Four labels at once: lanes 16q..16q+15 evaluate label q of
                // group k - kQuad for the chunk's 16 nodes, whose coefficients
                // come from the lanes holding those nodes.
                const int grp = k - kQuad;
                const int q = (t & 63) >> 4, n16 = t & 15;
                const int j = grp < 2 ? 4 * grp + q : (grp < 5 ? 4 * (grp - 2) + q : 2 * (grp - 5) + (q & 1));
                const int kc = grp < 2 ? kCornerK[j] : (grp < 5 ? kEdgeK[j] : kFaceK[j]);
                const int oy = kc / 9 - 1, ox = (kc / 3) % 3 - 1, oz = kc % 3 - 1;
                const int src = (m0 - z0) + n16;
                // Every lane evaluates (its tile indices stay in range for
                // any node of the segment) and only lanes whose neighbour
                // exists store, so the shuffles run with the whole wave.
                auto get = [&](int i) {
                    const cpx<R> a = i == 0 ? c.rx2 : i == 1 ? c.ry2 : i == 2 ? c.rz2 : i == 3 ? c.rxy :
                                     i == 4 ? c.rxz : i == 5 ? c.ryz : i == 6 ? c.rx1 : i == 7 ? c.ry1 : c.rz1;
                    return cpx<R>(__shfl(a.re, src), __shfl(a.im, src));
                };
                const R wmN = __shfl(grp < 2 ? c.wm4o8 : (grp < 5 ? c.wm3o12 : c.wm2o6), src);
                const int modeN = __shfl(mode, src);
                const int mN = m0 + n16, izN = mN + 1, tzN = mN - z0 + 1;
                R ro = s_rho[3 * (ox + 1) + (oy + 1)][tzN + oz];
                if (p.fst && izN == 1 && oz == -1) ro = R(0);
                const bool ok = (grp < 5 || q < 2) && mN < m1 && iy + oy >= 1 && iy + oy <= ny && ix + ox >= 1 &&
                                ix + ox <= nx && izN + oz >= 1 && izN + oz <= nz;
                const int cznN = 1 + (izN > 1) + (izN < nz), rlN = 3 * cxy * cznN;
                const int ry = oy > -1 ? (iy > 1) + (oy > 0) : 0, rx = ox > -1 ? (ix > 1) + (ox > 0) : 0;
                const int rz = oz > -1 ? (izN > 1) + (oz > 0) : 0;
                const int posN = 9 * cxy * (col_units(mN, nz) - col_units(m0, nz)) + ((ry * cxn + rx) * cznN + rz) * 3;
                auto emit = [&](int e, cpx<R> val) {
                    if (!ok) return;
                    // a group label is never the centre; traction-free nodes
                    // get their corners after the labels
                    if (modeN >= 2) val = cpx<R>(R(0), R(0));
                    V w;
                    w.x = val.re;
                    w.y = val.im;
                    stage[posN + (e / 3) * rlN + (e % 3)] = w;
                };
                if (grp < 2) {
                    corner_block<R>(kCorner[j], s_cell[tzN - 1], s_cell[tzN], get, wmN, ro, emit);
                } else if (grp == 2) {
                    edge_block<R, 0>(kEdge[j], s_cell[tzN - 1], s_cell[tzN], get, wmN, ro, emit);
                } else if (grp == 3) {
                    edge_block<R, 1>(kEdge[j], s_cell[tzN - 1], s_cell[tzN], get, wmN, ro, emit);
                } else if (grp == 4) {
                    edge_block<R, 2>(kEdge[j], s_cell[tzN - 1], s_cell[tzN], get, wmN, ro, emit);
                } else if (grp == 5) {
                    face_block<R, 0>(kFace[j], s_cell[tzN - 1], s_cell[tzN], get, wmN, ro, emit);
                } else if (grp == 6) {
                    face_block<R, 1>(kFace[j], s_cell[tzN - 1], s_cell[tzN], get, wmN, ro, emit);
                } else {
                    face_block<R, 2>(kFace[j], s_cell[tzN - 1], s_cell[tzN], get, wmN, ro, emit);
                }
                continue;
            }
            const int oy = k / 9 - 1, ox = (k / 3) % 3 - 1, oz = k % 3 - 1;
            if (iy + oy < 1 || iy + oy > ny || ix + ox < 1 || ix + ox > nx) continue;  // block-uniform
            if (!mine || iz + oz < 1 || iz + oz > nz) continue;
            // density the mass lumping uses at this neighbour (clamped gather,
            // top layer zeroed at the free surface, setcoefs3d.f:187-191)
            R ro = s_rho[3 * (ox + 1) + (oy + 1)][tz + oz];
            if (topfree && oz == -1) ro = R(0);
            cpx<R> v[9];
            label_block<R>(k, c, ro, v);
            if (mode >= 2) {
                // Dirichlet: identity centre block; traction-free: written after the labels
#pragma unroll
                for (int e = 0; e < 9; ++e)
                    v[e] = cpx<R>((mode == 2 && k == 13 && (e == 0 || e == 4 || e == 8)) ? R(1) : R(0), R(0));
            }
            const int ry = oy > -1 ? (iy > 1) + (oy > 0) : 0, rx = ox > -1 ? (ix > 1) + (ox > 0) : 0;
            const int rz = oz > -1 ? (iz > 1) + (oz > 0) : 0;
            const int pos = base + ((ry * cxn + rx) * czn + rz) * 3;
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int e = 0; e < 3; ++e) {
                    V w;
                    w.x = v[3 * r + e].re;
                    w.y = v[3 * r + e].im;
                    stage[pos + r * rl + e] = w;
                }
        }
        __syncthreads();
        if (p.fst && !p.usemin && m0 == 0) {  // block-uniform: the chunk holds the top node
            // the traction-free top node (lane 0 of wave 0 holds node iz = 1)
            // overwrites its five coupled blocks in the span
            if (t == 0 && mode == 3) tpfree_store<R>(c, ix, iy, nx, ny, nz, stage);
            __syncthreads();
        }
        // Span out.  The LDS reads of a half are issued before its first
        // store, so a wave waits on LDS twice per chunk; the stores drain
        // while the next chunk is evaluated.  They are nontemporal: the
        // values are not read again by this kernel, and streaming them past
        // the caches measured 11 % faster (0.638 -> 0.570 ms at 96x96x64).
        const int len = 9 * cxy * (col_units(m1, nz) - col_units(m0, nz));
        V* const dst = col + 9 * cxy * col_units(m0, nz);
        constexpr int NS = (STG_N + ASM_THREADS - 1) / ASM_THREADS;
        constexpr int NH = 4;
        typedef R nv2 __attribute__((ext_vector_type(2)));
#pragma unroll
        for (int h = 0; h < NS; h += NH) {
            V w[NH];
#pragma unroll
            for (int u = 0; u < NH; ++u) {
                const int i = t + (h + u) * ASM_THREADS;
                w[u] = stage[i < STG_N ? i : STG_N - 1];
            }
#pragma unroll
            for (int u = 0; u < NH; ++u)
                if (t + (h + u) * ASM_THREADS < len) {
                    const nv2 val = {w[u].x, w[u].y};
                    __builtin_nontemporal_store(val, (nv2*)&dst[t + (h + u) * ASM_THREADS]);
                }
        }
        __syncthreads();  // span read before the next chunk overwrites it
    }
}

// CSR pattern: row pointers and ascending column indices (base 0 or 1).
__global__ __launch_bounds__(ASM_WAVE) void assemble_pattern_kernel(
    int nx, int ny, int nz, int base, int32_t* __restrict__ rowptr, int32_t* __restrict__ colidx)
{
    const int iz = blockIdx.x * ASM_WAVE + threadIdx.x + 1;
    const int ix = blockIdx.y + 1, iy = blockIdx.z + 1;
    if (iz > nz) return;
    const int64_t g0 = (int64_t)(iz - 1) + (int64_t)(ix - 1) * nz + (int64_t)(iy - 1) * nx * nz;
    const int cxn = 1 + (ix > 1) + (ix < nx), cyn = 1 + (iy > 1) + (iy < ny), czn = 1 + (iz > 1) + (iz < nz);
    const int rowlen = 3 * cxn * cyn * czn;
    const int64_t off = node_offset(iz, ix, iy, nx, ny, nz);
#pragma unroll
    for (int r = 0; r < 3; ++r) rowptr[3 * g0 + r] = (int32_t)(off + r * rowlen + base);
    if (ix == nx && iy == ny && iz == nz) rowptr[3 * g0 + 3] = (int32_t)(off + 3 * rowlen + base);
    int pos = 0;
    for (int oy = -1; oy <= 1; ++oy) {
        if (iy + oy < 1 || iy + oy > ny) continue;
        for (int ox = -1; ox <= 1; ++ox) {
            if (ix + ox < 1 || ix + ox > nx) continue;
            for (int oz = -1; oz <= 1; ++oz) {
                if (iz + oz < 1 || iz + oz > nz) continue;
                const int64_t gn = g0 + oz + (int64_t)ox * nz + (int64_t)oy * nx * nz;
#pragma unroll
                for (int r = 0; r < 3; ++r) {
                    int32_t* dst = colidx + off + r * rowlen + pos;
                    dst[0] = (int32_t)(3 * gn + base);
                    dst[1] = (int32_t)(3 * gn + 1 + base);
                    dst[2] = (int32_t)(3 * gn + 2 + base);
                }
                pos += 3;
            }
        }
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Host launchers (C++ linkage; the C ABI lives in wt_api.cpp)

int64_t wt_nnz(int nx, int ny, int nz)
{
    return 9 * axis_prefix(nx + 1, nx) * axis_prefix(ny + 1, ny) * axis_prefix(nz + 1, nz);
}

int wt_launch_pattern(int nx, int ny, int nz, int base, int32_t* d_rowptr, int32_t* d_colidx,
                      hipStream_t stream)
{
    if (nx < 1 || ny < 1 || nz < 1) return WT_ERR_ARG;
    dim3 grid((nz + ASM_WAVE - 1) / ASM_WAVE, nx, ny);
    hipLaunchKernelGGL(assemble_pattern_kernel, grid, dim3(ASM_WAVE), 0, stream, nx, ny, nz, base,
                       d_rowptr, d_colidx);
    WT_HIP(hipGetLastError());
    return WT_OK;
}

template <typename R>
int wt_launch_values(const wt_grid_params& p, const void* d_lam, const void* d_mu, const void* d_rho,
                     void* d_vals, int* d_err, hipStream_t stream)
{
    if (p.nx < 2 || p.ny < 2 || p.nz < 2) return WT_ERR_ARG;
    const int64_t nblk = (int64_t)((p.nz + ASM_WAVE - 1) / ASM_WAVE) * p.nx * p.ny;
    if (nblk > 0x7fffffff) return WT_ERR_ARG;
    hipLaunchKernelGGL(assemble_values_kernel<R>, dim3((unsigned)nblk), dim3(ASM_THREADS), 0, stream, p,
                       (const cpx<R>*)d_lam, (const cpx<R>*)d_mu, (const R*)d_rho, (cpx<R>*)d_vals, d_err,
                       nblk);
    WT_HIP(hipGetLastError());
    return WT_OK;
}

template int wt_launch_values<float>(const wt_grid_params&, const void*, const void*, const void*, void*,
                                     int*, hipStream_t);
template int wt_launch_values<double>(const wt_grid_params&, const void*, const void*, const void*, void*,
                                      int*, hipStream_t);
