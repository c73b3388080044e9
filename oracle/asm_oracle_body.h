/* Precision-generic body of the assembly oracle; included by asm_oracle.c
 * once with R=float (SFX f32) and once with R=double (SFX f64).
 * TEST INFRASTRUCTURE ONLY: see the header of asm_oracle.c. */

#define CT CAT(cpx_, SFX)
#define FN(name) CAT(name, SFX)

typedef struct { R re, im; } CT;

static inline CT FN(mk_)(R re, R im) { CT z; z.re = re; z.im = im; return z; }
static inline CT FN(add_)(CT a, CT b) { return FN(mk_)(a.re + b.re, a.im + b.im); }
static inline CT FN(sub_)(CT a, CT b) { return FN(mk_)(a.re - b.re, a.im - b.im); }
static inline CT FN(neg_)(CT a) { return FN(mk_)(-a.re, -a.im); }
static inline CT FN(mul_)(CT a, CT b)
{
    R re = a.re * b.re; re = re - a.im * b.im;
    R im = a.re * b.im; im = im + a.im * b.re;
    return FN(mk_)(re, im);
}
static inline CT FN(scl_)(R s, CT a) { return FN(mk_)(s * a.re, s * a.im); }
static inline CT FN(divr_)(CT a, R s) { return FN(mk_)(a.re / s, a.im / s); }
/* Smith's complex division */
static inline CT FN(div_)(CT a, CT b)
{
    R ar = a.re, ai = a.im, br = b.re, bi = b.im;
    if (FABS(br) >= FABS(bi)) {
        R t = bi / br; R den = br + bi * t;
        return FN(mk_)((ar + ai * t) / den, (ai - ar * t) / den);
    } else {
        R t = br / bi; R den = bi + br * t;
        return FN(mk_)((ar * t + ai) / den, (ai * t - ar) / den);
    }
}
/* real part of the principal square root */
static inline R FN(csqrt_re_)(CT z)
{
    if (z.im == (R)0) return z.re >= (R)0 ? SQRT(z.re) : (R)0;
    R m = HYPOT(z.re, z.im);
    return SQRT((m + z.re) / (R)2);
}

/* mass-lumping weights, calcwts.f:28-71 (x = vp/vs is formed in R) */
static void FN(calcwts_)(R vp, R vs, R *wm)
{
    static const double p1[4] = {0.2909366e+00, -0.1792402e-01, 0.7329184e-02, -0.1055034e-02};
    static const double p2[4] = {0.6367377e+00, -0.7982022e-02, 0.3385164e-02, -0.5065399e-03};
    static const double p3[4] = {0.7244875e-01,  0.2569959e-01, -0.1060011e-01, 0.1540534e-02};
    double x = (double)(vp / vs);
    double w2 = 0.0, w3 = 0.0, w4;
    double w1 = p1[0] + x * (p1[1] + x * (p1[2] + x * p1[3]));
    if (w1 > 1.0) w1 = 1.0;
    else w2 = p2[0] + x * (p2[1] + x * (p2[2] + x * p2[3]));
    if ((w1 + w2) > 1.0) w2 = 0.0;
    else w3 = p3[0] + x * (p3[1] + x * (p3[2] + x * p3[3]));
    if ((w1 + w2 + w3) > 1.0) w3 = 1.0 - (w1 + w2);
    w4 = 1.0 - (w1 + w2 + w3);
    wm[0] = (R)w1; wm[1] = (R)w2; wm[2] = (R)w3; wm[3] = (R)w4;
}

typedef struct {
    int nx, ny, nz, ipml, fst, usemin;
    R dx, pmld, pmlf;
    CT omega;
} FN(prm_);

/* per-axis PML stretch factors, pml3d.f:207-295 (kmax = 1, fdom = 0) */
static int FN(pml_axis_)(const FN(prm_) *P, int i, int n, R vpreal, CT eyeom,
                         CT *r1, CT *r2)
{
    R dpml = 0; int isn = 0;
    if (i <= P->ipml) { dpml = P->pmld - (R)(i - 1) * P->dx; isn = 1; }
    else if ((n - i + 1) <= P->ipml) { dpml = P->pmld - (R)(n - i) * P->dx; isn = -1; }
    if (dpml < 0) return 1;
    R dn = dpml * dpml; dn = dn * vpreal; dn = dn * P->pmlf;
    CT num = eyeom;                                  /* alpha + (-i w), alpha = 0 */
    CT den = FN(mk_)(eyeom.re + dn, eyeom.im);        /* kappa = 1 */
    *r1 = FN(div_)(num, den);
    CT r1sq = FN(mul_)(*r1, *r1);
    CT t = FN(div_)(eyeom, FN(mul_)(eyeom, eyeom));
    R s = (R)isn * (R)2;
    CT q = FN(scl_)(s, r1sq);
    q = FN(mul_)(q, *r1);
    q = FN(scl_)(dpml, q);
    R vf = vpreal * P->pmlf;
    CT inner = FN(scl_)(vf, t);
    *r2 = FN(mul_)(q, inner);
    return 0;
}

/* All 27 3x3 blocks of one grid node in the stretched-coordinate form of
 * pml3d.f:344-800.  The interior form pfdfd3d.f:162-576 is the same
 * expression with r1 = 1, r2 = 0 (fdfd3d.f:56-57 notes this identity).
 * blk[k][3*r+c], k = 9(oy+1)+3(ox+1)+(oz+1), r,c in {u,v,w}.
 * mode: 0 interior, 1 PML, 2 Dirichlet.  izp: z index used for the PML. */
static int FN(node_blocks_)(const FN(prm_) *P, const CT *lam, const CT *mu,
                            const R *rho, int iz, int ix, int iy,
                            int mode, int izp, CT blk[27][9])
{
    int nx = P->nx, ny = P->ny, nz = P->nz;
    CT da[3][3][3], mm[3][3][3];
    R ro[3][3][3];
    if (mode == 2) {
        for (int k = 0; k < 27; ++k)
            for (int e = 0; e < 9; ++e) blk[k][e] = FN(mk_)(0, 0);
        blk[13][0] = FN(mk_)(1, 0); blk[13][4] = FN(mk_)(1, 0); blk[13][8] = FN(mk_)(1, 0);
        return 0;
    }
    /* setcoefs3d.f:79-105: clamped 3x3x3 gather, (i=x, j=y, k=z) */
    for (int k = 0; k < 3; ++k) {
        int zg = iz + k - 1; if (zg < 1) zg = 1; if (zg > nz) zg = nz;
        for (int i = 0; i < 3; ++i) {
            int xg = ix + i - 1; if (xg < 1) xg = 1; if (xg > nx) xg = nx;
            for (int j = 0; j < 3; ++j) {
                int yg = iy + j - 1; if (yg < 1) yg = 1; if (yg > ny) yg = ny;
                long g = (long)(zg - 1) + (long)(xg - 1) * nz + (long)(yg - 1) * nx * nz;
                da[i][j][k] = lam[g]; mm[i][j][k] = mu[g]; ro[i][j][k] = rho[g];
            }
        }
    }
    /* setcoefs3d.f:108-174: 8 octant cells */
    CT u[9], d[9];
    for (int pass = 0; pass < 2; ++pass) {
        CT (*a)[3][3] = pass ? da : mm;
        CT *o = pass ? d : u;
        CT km[3][4];
        for (int k = 0; k < 3; ++k) {
            km[k][0] = FN(divr_)(FN(add_)(FN(add_)(FN(add_)(a[0][0][k], a[1][0][k]), a[0][1][k]), a[1][1][k]), (R)4);
            km[k][1] = FN(divr_)(FN(add_)(FN(add_)(FN(add_)(a[0][2][k], a[1][2][k]), a[0][1][k]), a[1][1][k]), (R)4);
            km[k][2] = FN(divr_)(FN(add_)(FN(add_)(FN(add_)(a[2][0][k], a[1][0][k]), a[2][1][k]), a[1][1][k]), (R)4);
            km[k][3] = FN(divr_)(FN(add_)(FN(add_)(FN(add_)(a[2][2][k], a[1][2][k]), a[2][1][k]), a[1][1][k]), (R)4);
        }
        for (int c = 0; c < 4; ++c) {
            o[1 + c] = FN(divr_)(FN(add_)(km[0][c], km[1][c]), (R)2);
            o[5 + c] = FN(divr_)(FN(add_)(km[2][c], km[1][c]), (R)2);
        }
    }
    /* setcoefs3d.f:177-192: free surface */
    if (P->fst && iz == 1) {
        for (int c = 1; c <= 4; ++c) { u[c] = FN(mk_)(0, 0); d[c] = FN(mk_)(0, 0); }
        for (int j = 0; j < 3; ++j) for (int i = 0; i < 3; ++i) ro[i][j][0] = 0;
    }
    CT lmu[9], lpu[9], lp2u[9];
    for (int c = 1; c <= 8; ++c) {
        lmu[c] = FN(sub_)(d[c], u[c]);
        lpu[c] = FN(add_)(d[c], u[c]);
        lp2u[c] = FN(add_)(lpu[c], u[c]);
    }
    /* setcoefs3d.f:224-225 + calcwts */
    R vp = FN(csqrt_re_)(FN(divr_)(FN(add_)(da[1][1][1], FN(scl_)((R)2, mm[1][1][1])), ro[1][1][1]));
    R vs = FN(csqrt_re_)(FN(divr_)(mm[1][1][1], ro[1][1][1]));
    R wm[4];
    FN(calcwts_)(vp, vs, wm);

    /* pml3d.f:176-198 / pfdfd3d.f:133-160 */
    CT om = P->omega;
    CT eyeom = FN(mk_)(om.im, -om.re);
    CT osq = FN(mul_)(om, om);
    R wm2o6 = (osq.re * wm[1]) / (R)6;
    R wm3o12 = (osq.re * wm[2]) / (R)12;
    R wm4o8 = (osq.re * wm[3]) / (R)8;
    R dx = P->dx, dx2 = dx + dx, dxx = dx * dx, dxx2 = dxx + dxx, dxx4 = dxx2 + dxx2;
    R vpreal = vp;
    CT one = FN(mk_)(1, 0), zero = FN(mk_)(0, 0);
    CT r1x = one, r1y = one, r1z = one, r2x = zero, r2y = zero, r2z = zero;
    if (mode == 3) {
        /* tpfree.f:55-420: the stretch factors r1 of the x and y PML (kmax = 1,
         * fdom = 0, the same expression as pml3d.f's r1), r1z = 1, no mass
         * term; only the z+1 corners (cells 5-8 below the node) and the centre
         * couple.  The centre's VW entry uses the lambda cells where UW uses
         * mu, as tpfree.f:321 has it. */
        if (FN(pml_axis_)(P, ix, nx, vpreal, eyeom, &r1x, &r2x)) return 1;
        if (FN(pml_axis_)(P, iy, ny, vpreal, eyeom, &r1y, &r2y)) return 1;
        R dxo8 = (R)1 / (dx * (R)8);
        CT X = FN(scl_)(dxo8, r1x), Y = FN(scl_)(dxo8, r1y), Z = FN(scl_)(dxo8, r1z);
        for (int k = 0; k < 27; ++k)
            for (int e = 0; e < 9; ++e) blk[k][e] = zero;
        /* corners: label index, cell, signs of the x and y terms */
        static const int ck[4] = {2, 20, 8, 26}, cc[4] = {5, 6, 7, 8}, sx[4] = {-1, -1, 1, 1}, sy[4] = {-1, 1, -1, 1};
        for (int q = 0; q < 4; ++q) {
            int c = cc[q];
            CT *b = blk[ck[q]];
            CT xu = FN(mul_)(X, u[c]), yu = FN(mul_)(Y, u[c]), xd = FN(mul_)(X, d[c]), yd = FN(mul_)(Y, d[c]);
            b[0] = FN(mul_)(Z, u[c]);                      /* UU */
            b[2] = sx[q] > 0 ? xu : FN(neg_)(xu);          /* UW */
            b[4] = FN(mul_)(Z, u[c]);                      /* VV */
            b[5] = sy[q] > 0 ? yu : FN(neg_)(yu);          /* VW */
            b[6] = sx[q] > 0 ? xd : FN(neg_)(xd);          /* WU */
            b[7] = sy[q] > 0 ? yd : FN(neg_)(yd);          /* WV */
            b[8] = FN(mul_)(Z, lp2u[c]);                   /* WW */
        }
        CT *b = blk[13];
        CT su = FN(add_)(FN(add_)(FN(add_)(u[8], u[7]), u[6]), u[5]);
        CT sl = FN(add_)(FN(add_)(FN(add_)(lp2u[8], lp2u[7]), lp2u[6]), lp2u[5]);
        CT dy = FN(add_)(FN(sub_)(FN(sub_)(d[5], d[8]), d[6]), d[7]);
        b[0] = FN(neg_)(FN(mul_)(Z, su));
        b[2] = FN(mul_)(X, FN(sub_)(FN(add_)(FN(sub_)(u[5], u[8]), u[6]), u[7]));
        b[4] = FN(neg_)(FN(mul_)(Z, su));
        b[5] = FN(mul_)(Y, dy);
        b[6] = FN(mul_)(X, FN(sub_)(FN(add_)(FN(sub_)(d[5], d[8]), d[6]), d[7]));
        b[7] = FN(mul_)(Y, dy);
        b[8] = FN(neg_)(FN(mul_)(Z, sl));
        return 0;
    }
    if (mode == 1) {
        if (FN(pml_axis_)(P, ix, nx, vpreal, eyeom, &r1x, &r2x)) return 1;
        if (FN(pml_axis_)(P, iy, ny, vpreal, eyeom, &r1y, &r2y)) return 1;
        if (FN(pml_axis_)(P, izp, nz, vpreal, eyeom, &r1z, &r2z)) return 1;
    }
    CT rx1 = FN(divr_)(FN(divr_)(r2x, dx2), (R)4);
    CT ry1 = FN(divr_)(FN(divr_)(r2y, dx2), (R)4);
    CT rz1 = FN(divr_)(FN(divr_)(r2z, dx2), (R)4);
    CT qx = FN(divr_)(FN(divr_)(FN(mul_)(r1x, r1x), dxx), (R)4);
    CT qy = FN(divr_)(FN(divr_)(FN(mul_)(r1y, r1y), dxx), (R)4);
    CT qz = FN(divr_)(FN(divr_)(FN(mul_)(r1z, r1z), dxx), (R)4);
    CT rxy = FN(divr_)(FN(divr_)(FN(mul_)(r1x, r1y), dxx4), (R)4);
    CT rxz = FN(divr_)(FN(divr_)(FN(mul_)(r1x, r1z), dxx4), (R)4);
    CT ryz = FN(divr_)(FN(divr_)(FN(mul_)(r1y, r1z), dxx4), (R)4);
    CT rx2 = FN(divr_)(qx, (R)2), ry2 = FN(divr_)(qy, (R)2), rz2 = FN(divr_)(qz, (R)2);
    CT r2xpy = FN(add_)(rx1, ry1), r2xmy = FN(sub_)(rx1, ry1);
    CT r2xpz = FN(add_)(rx1, rz1), r2xmz = FN(sub_)(rx1, rz1);
    CT r2ypz = FN(add_)(ry1, rz1), r2ymz = FN(sub_)(ry1, rz1);
    CT rxpy = FN(add_)(rx2, ry2), rxpz = FN(add_)(rx2, rz2), rypz = FN(add_)(ry2, rz2);

#define M_(a, b) FN(mul_)(a, b)
#define A_(a, b) FN(add_)(a, b)
#define S_(a, b) FN(sub_)(a, b)
#define N_(a) FN(neg_)(a)
#define MS(k) FN(mk_)(k * ro[ox][oy][oz], 0)
    /* corner-type diagonal: q1*lp2u + qs*u (+/-) r2a*lp2u (+/-) r2b*u + massf */
#define DG(q1, qs, sa, r2a, sb, r2b, L, U, mf) \
    A_(A_(A_(A_(M_(q1, L), M_(qs, U)), sa > 0 ? M_(r2a, L) : N_(M_(r2a, L))), \
          sb > 0 ? M_(r2b, U) : N_(M_(r2b, U))), mf)
#define SET9(kk, uu, vv, ww, uv, vu, uw, wu, vw, wv) do { CT *b_ = blk[kk]; \
    b_[0] = uu; b_[1] = uv; b_[2] = uw; b_[3] = vu; b_[4] = vv; b_[5] = vw; \
    b_[6] = wu; b_[7] = wv; b_[8] = ww; } while (0)
    int ox, oy, oz;
    CT mf, t0, t1, t2, t3, t4, t5;
    /* ---- oy = -1 (m) ---- */
    oy = 0;
    ox = 0; oz = 0; mf = MS(wm4o8);   /* a1m adm */
    t0 = M_(rxy, lpu[1]); t1 = M_(rxz, lpu[1]); t2 = M_(ryz, lpu[1]);
    SET9(0, DG(rx2, rypz, -1, rx1, -1, r2ypz, lp2u[1], u[1], mf),
            DG(ry2, rxpz, -1, ry1, -1, r2xpz, lp2u[1], u[1], mf),
            DG(rz2, rxpy, -1, rz1, -1, r2xpy, lp2u[1], u[1], mf),
            t0, t0, t1, t1, t2, t2);
    ox = 0; oz = 1; mf = MS(wm3o12);  /* a2m aam */
    t0 = A_(N_(M_(rz2, A_(u[1], u[5]))), mf); t1 = A_(N_(M_(rz2, A_(lp2u[1], lp2u[5]))), mf);
    t2 = M_(rxy, A_(lpu[1], lpu[5])); t3 = N_(M_(rxz, S_(lmu[1], lmu[5]))); t4 = N_(M_(ryz, S_(lmu[1], lmu[5])));
    SET9(1, t0, t0, t1, t2, t2, t3, N_(t3), t4, N_(t4));
    ox = 0; oz = 2; mf = MS(wm4o8);   /* a3m afm */
    t0 = M_(rxy, lpu[5]); t1 = N_(M_(rxz, lpu[5])); t2 = N_(M_(ryz, lpu[5]));
    SET9(2, DG(rx2, rypz, -1, rx1, -1, r2ymz, lp2u[5], u[5], mf),
            DG(ry2, rxpz, -1, ry1, -1, r2xmz, lp2u[5], u[5], mf),
            DG(rz2, rxpy, +1, rz1, -1, r2xmy, lp2u[5], u[5], mf),
            t0, t0, t1, t1, t2, t2);
    ox = 1; oz = 0; mf = MS(wm3o12);  /* a4m ddm */
    t0 = A_(N_(M_(rx2, A_(lp2u[1], lp2u[3]))), mf); t1 = A_(N_(M_(rx2, A_(u[1], u[3]))), mf);
    t2 = M_(rxy, S_(lmu[1], lmu[3])); t3 = M_(rxz, S_(lmu[1], lmu[3])); t4 = M_(ryz, A_(lpu[1], lpu[3]));
    SET9(3, t0, t0, t1, t2, N_(t2), t3, N_(t3), t4, t4);
    ox = 1; oz = 1; mf = MS(wm2o6);   /* a5m bem */
    t0 = A_(M_(ry2, A_(A_(A_(u[1], u[7]), u[5]), u[3])), mf);
    t2 = M_(rxy, S_(A_(S_(lmu[1], lmu[7]), lmu[5]), lmu[3]));
    t3 = N_(M_(rxz, S_(S_(A_(lpu[1], lpu[7]), lpu[5]), lpu[3])));
    t4 = N_(M_(ryz, A_(S_(S_(lmu[1], lmu[7]), lmu[5]), lmu[3])));
    SET9(4, t0, t0, t0, t2, N_(t2), t3, t3, t4, N_(t4));
    ox = 1; oz = 2; mf = MS(wm3o12);  /* a6m ffm */
    t0 = A_(N_(M_(rx2, A_(lp2u[7], lp2u[5]))), mf); t1 = A_(N_(M_(rx2, A_(u[7], u[5]))), mf);
    t2 = M_(rxy, S_(lmu[5], lmu[7])); t3 = M_(rxz, S_(lmu[7], lmu[5])); t4 = N_(M_(ryz, A_(lpu[7], lpu[5])));
    SET9(5, t0, t0, t1, t2, N_(t2), t3, N_(t3), t4, t4);
    ox = 2; oz = 0; mf = MS(wm4o8);   /* a7m cdm */
    t0 = N_(M_(rxy, lpu[3])); t1 = N_(M_(rxz, lpu[3])); t2 = M_(ryz, lpu[3]);
    SET9(6, DG(rx2, rypz, +1, rx1, -1, r2ypz, lp2u[3], u[3], mf),
            DG(ry2, rxpz, -1, ry1, +1, r2xmz, lp2u[3], u[3], mf),
            DG(rz2, rxpy, -1, rz1, +1, r2xpy, lp2u[3], u[3], mf),
            t0, t0, t1, t1, t2, t2);
    ox = 2; oz = 1; mf = MS(wm3o12);  /* a8m ccm */
    t0 = A_(N_(M_(rz2, A_(u[7], u[3]))), mf); t1 = A_(N_(M_(rz2, A_(lp2u[7], lp2u[3]))), mf);
    t2 = N_(M_(rxy, A_(lpu[7], lpu[3]))); t3 = N_(M_(rxz, S_(lmu[7], lmu[3]))); t4 = N_(M_(ryz, S_(lmu[3], lmu[7])));
    SET9(7, t0, t0, t1, t2, t2, t3, N_(t3), t4, N_(t4));
    ox = 2; oz = 2; mf = MS(wm4o8);   /* a9m cfm */
    t0 = N_(M_(rxy, lpu[7])); t1 = M_(rxz, lpu[7]); t2 = N_(M_(ryz, lpu[7]));
    SET9(8, DG(rx2, rypz, +1, rx1, -1, r2ymz, lp2u[7], u[7], mf),
            DG(ry2, rxpz, -1, ry1, +1, r2xpz, lp2u[7], u[7], mf),
            DG(rz2, rxpy, +1, rz1, +1, r2xmy, lp2u[7], u[7], mf),
            t0, t0, t1, t1, t2, t2);
    /* ---- oy = 0 (n) ---- */
    oy = 1;
    ox = 0; oz = 0; mf = MS(wm3o12);  /* a1n adn */
    t0 = A_(N_(M_(ry2, A_(u[1], u[2]))), mf);
    t2 = M_(rxy, S_(lmu[2], lmu[1])); t3 = M_(rxz, A_(lpu[1], lpu[2])); t4 = M_(ryz, S_(lmu[1], lmu[2]));
    SET9(9, t0, t0, t0, t2, N_(t2), t3, t3, t4, N_(t4));
    ox = 0; oz = 1; mf = MS(wm2o6);   /* a2n aan */
    t0 = A_(M_(rx2, A_(A_(A_(lp2u[1], lp2u[6]), lp2u[2]), lp2u[5])), mf);
    t1 = A_(M_(rx2, A_(A_(A_(u[1], u[6]), u[2]), u[5])), mf);
    t2 = M_(rxy, S_(A_(S_(lmu[2], lmu[5]), lmu[6]), lmu[1]));
    t3 = N_(M_(rxz, S_(A_(S_(lmu[2], lmu[5]), lmu[1]), lmu[6])));
    t4 = M_(ryz, S_(S_(A_(lpu[2], lpu[5]), lpu[1]), lpu[6]));
    SET9(10, t0, t0, t1, t2, N_(t2), t3, N_(t3), t4, t4);
    ox = 0; oz = 2; mf = MS(wm3o12);  /* a3n afn */
    t0 = A_(N_(M_(ry2, A_(u[5], u[6]))), mf);
    t2 = N_(M_(rxy, S_(lmu[5], lmu[6]))); t3 = N_(M_(rxz, A_(lpu[5], lpu[6]))); t4 = N_(M_(ryz, S_(lmu[5], lmu[6])));
    SET9(11, t0, t0, t0, t2, N_(t2), t3, t3, t4, N_(t4));
    ox = 1; oz = 0; mf = MS(wm2o6);   /* a4n ddn */
    t0 = A_(M_(rz2, A_(A_(A_(u[1], u[4]), u[2]), u[3])), mf);
    t1 = A_(M_(rz2, A_(A_(A_(lp2u[1], lp2u[4]), lp2u[2]), lp2u[3])), mf);
    t2 = N_(M_(rxy, S_(S_(A_(lpu[1], lpu[4]), lpu[2]), lpu[3])));
    t3 = M_(rxz, S_(A_(S_(lmu[1], lmu[4]), lmu[2]), lmu[3]));
    t4 = M_(ryz, S_(A_(S_(lmu[1], lmu[4]), lmu[3]), lmu[2]));
    SET9(12, t0, t0, t1, t2, t2, t3, N_(t3), t4, N_(t4));
    ox = 1; oz = 1;                   /* a5n ben */
    {
        CT mc = FN(scl_)(wm[0], osq); mc = FN(mk_)(mc.re * ro[1][1][1], mc.im * ro[1][1][1]);
        CT ls = lp2u[1], us = u[1];
        for (int c = 2; c <= 8; ++c) { ls = A_(ls, lp2u[c]); us = A_(us, u[c]); }
        CT ex, ey, ez, fa, fb;
        ex = S_(A_(S_(A_(S_(A_(S_(lp2u[1], lp2u[8]), lp2u[2]), lp2u[7]), lp2u[5]), lp2u[4]), lp2u[6]), lp2u[3]);
        fa = S_(A_(S_(u[1], u[8]), u[3]), u[6]);
        fb = S_(A_(S_(u[7], u[2]), u[5]), u[4]);
        t0 = A_(A_(A_(A_(S_(N_(M_(rx2, ls)), M_(rypz, us)), M_(rx1, ex)), M_(r2ypz, fa)), M_(r2ymz, fb)), mc);
        ey = S_(A_(S_(A_(S_(A_(S_(lp2u[1], lp2u[2]), lp2u[5]), lp2u[6]), lp2u[7]), lp2u[8]), lp2u[3]), lp2u[4]);
        fa = S_(A_(S_(u[1], u[8]), u[2]), u[7]);
        fb = S_(A_(S_(u[6], u[3]), u[5]), u[4]);
        t1 = A_(A_(A_(A_(S_(N_(M_(ry2, ls)), M_(rxpz, us)), M_(ry1, ey)), M_(r2xpz, fa)), M_(r2xmz, fb)), mc);
        ez = S_(A_(S_(A_(S_(A_(S_(lp2u[1], lp2u[8]), lp2u[2]), lp2u[7]), lp2u[4]), lp2u[5]), lp2u[3]), lp2u[6]);
        fa = S_(A_(S_(u[1], u[8]), u[5]), u[4]);
        fb = S_(A_(S_(u[2], u[7]), u[6]), u[3]);
        t2 = A_(A_(A_(A_(S_(N_(M_(rz2, ls)), M_(rxpy, us)), M_(rz1, ez)), M_(r2xpy, fa)), M_(r2xmy, fb)), mc);
        t3 = M_(rxy, S_(A_(S_(A_(S_(A_(S_(lpu[2], lpu[1]), lpu[7]), lpu[8]), lpu[3]), lpu[4]), lpu[6]), lpu[5]));
        t4 = M_(rxz, S_(A_(S_(A_(S_(A_(S_(lpu[4], lpu[1]), lpu[5]), lpu[8]), lpu[3]), lpu[2]), lpu[6]), lpu[7]));
        t5 = M_(ryz, S_(A_(S_(A_(S_(A_(S_(lpu[2], lpu[1]), lpu[7]), lpu[8]), lpu[4]), lpu[3]), lpu[5]), lpu[6]));
        SET9(13, t0, t1, t2, t3, t3, t4, t4, t5, t5);
    }
    ox = 1; oz = 2; mf = MS(wm2o6);   /* a6n ffn */
    t0 = A_(M_(rz2, A_(A_(A_(u[8], u[5]), u[7]), u[6])), mf);
    t1 = A_(M_(rz2, A_(A_(A_(lp2u[8], lp2u[5]), lp2u[7]), lp2u[6])), mf);
    t2 = N_(M_(rxy, S_(S_(A_(lpu[8], lpu[5]), lpu[7]), lpu[6])));
    t3 = M_(rxz, S_(A_(S_(lmu[8], lmu[5]), lmu[7]), lmu[6]));
    t4 = M_(ryz, S_(A_(S_(lmu[8], lmu[5]), lmu[6]), lmu[7]));
    SET9(14, t0, t0, t1, t2, t2, t3, N_(t3), t4, N_(t4));
    ox = 2; oz = 0; mf = MS(wm3o12);  /* a7n cdn */
    t0 = A_(N_(M_(ry2, A_(u[4], u[3]))), mf);
    t2 = N_(M_(rxy, S_(lmu[4], lmu[3]))); t3 = N_(M_(rxz, A_(lpu[4], lpu[3]))); t4 = N_(M_(ryz, S_(lmu[4], lmu[3])));
    SET9(15, t0, t0, t0, t2, N_(t2), t3, t3, t4, N_(t4));
    ox = 2; oz = 1; mf = MS(wm2o6);   /* a8n ccn */
    t0 = A_(M_(rx2, A_(A_(A_(lp2u[8], lp2u[3]), lp2u[7]), lp2u[4])), mf);
    t1 = A_(M_(rx2, A_(A_(A_(u[8], u[3]), u[7]), u[4])), mf);
    t2 = M_(rxy, S_(A_(S_(lmu[7], lmu[4]), lmu[3]), lmu[8]));
    t3 = N_(M_(rxz, S_(A_(S_(lmu[7], lmu[4]), lmu[8]), lmu[3])));
    t4 = M_(ryz, S_(S_(A_(lpu[7], lpu[4]), lpu[8]), lpu[3]));
    SET9(16, t0, t0, t1, t2, N_(t2), t3, N_(t3), t4, t4);
    ox = 2; oz = 2; mf = MS(wm3o12);  /* a9n cfn */
    t0 = A_(N_(M_(ry2, A_(u[8], u[7]))), mf);
    t2 = M_(rxy, S_(lmu[7], lmu[8])); t3 = M_(rxz, A_(lpu[8], lpu[7])); t4 = M_(ryz, S_(lmu[8], lmu[7]));
    SET9(17, t0, t0, t0, t2, N_(t2), t3, t3, t4, N_(t4));
    /* ---- oy = +1 (p) ---- */
    oy = 2;
    ox = 0; oz = 0; mf = MS(wm4o8);   /* a1p adp */
    t0 = N_(M_(rxy, lpu[2])); t1 = M_(rxz, lpu[2]); t2 = N_(M_(ryz, lpu[2]));
    SET9(18, DG(rx2, rypz, -1, rx1, +1, r2ymz, lp2u[2], u[2], mf),
             DG(ry2, rxpz, +1, ry1, -1, r2xpz, lp2u[2], u[2], mf),
             DG(rz2, rxpy, -1, rz1, -1, r2xmy, lp2u[2], u[2], mf),
             t0, t0, t1, t1, t2, t2);
    ox = 0; oz = 1; mf = MS(wm3o12);  /* a2p aap */
    t0 = A_(N_(M_(rz2, A_(u[2], u[6]))), mf); t1 = A_(N_(M_(rz2, A_(lp2u[2], lp2u[6]))), mf);
    t2 = N_(M_(rxy, A_(lpu[2], lpu[6]))); t3 = N_(M_(rxz, S_(lmu[2], lmu[6]))); t4 = N_(M_(ryz, S_(lmu[6], lmu[2])));
    SET9(19, t0, t0, t1, t2, t2, t3, N_(t3), t4, N_(t4));
    ox = 0; oz = 2; mf = MS(wm4o8);   /* a3p afp */
    t0 = N_(M_(rxy, lpu[6])); t1 = N_(M_(rxz, lpu[6])); t2 = M_(ryz, lpu[6]);
    SET9(20, DG(rx2, rypz, -1, rx1, +1, r2ypz, lp2u[6], u[6], mf),
             DG(ry2, rxpz, +1, ry1, -1, r2xmz, lp2u[6], u[6], mf),
             DG(rz2, rxpy, +1, rz1, -1, r2xpy, lp2u[6], u[6], mf),
             t0, t0, t1, t1, t2, t2);
    ox = 1; oz = 0; mf = MS(wm3o12);  /* a4p ddp */
    t0 = A_(N_(M_(rx2, A_(lp2u[2], lp2u[4]))), mf); t1 = A_(N_(M_(rx2, A_(u[2], u[4]))), mf);
    t2 = M_(rxy, S_(lmu[4], lmu[2])); t3 = M_(rxz, S_(lmu[2], lmu[4])); t4 = N_(M_(ryz, A_(lpu[2], lpu[4])));
    SET9(21, t0, t0, t1, t2, N_(t2), t3, N_(t3), t4, t4);
    ox = 1; oz = 1; mf = MS(wm2o6);   /* a5p bep */
    t0 = A_(M_(ry2, A_(A_(A_(u[8], u[2]), u[4]), u[6])), mf);
    t2 = M_(rxy, S_(A_(S_(lmu[8], lmu[2]), lmu[4]), lmu[6]));
    t3 = N_(M_(rxz, S_(S_(A_(lpu[8], lpu[2]), lpu[4]), lpu[6])));
    t4 = N_(M_(ryz, A_(S_(S_(lmu[8], lmu[2]), lmu[4]), lmu[6])));
    SET9(22, t0, t0, t0, t2, N_(t2), t3, t3, t4, N_(t4));
    ox = 1; oz = 2; mf = MS(wm3o12);  /* a6p ffp */
    t0 = A_(N_(M_(rx2, A_(lp2u[8], lp2u[6]))), mf); t1 = A_(N_(M_(rx2, A_(u[8], u[6]))), mf);
    t2 = M_(rxy, S_(lmu[8], lmu[6])); t3 = M_(rxz, S_(lmu[8], lmu[6])); t4 = M_(ryz, A_(lpu[8], lpu[6]));
    SET9(23, t0, t0, t1, t2, N_(t2), t3, N_(t3), t4, t4);
    ox = 2; oz = 0; mf = MS(wm4o8);   /* a7p cdp */
    t0 = M_(rxy, lpu[4]); t1 = N_(M_(rxz, lpu[4])); t2 = N_(M_(ryz, lpu[4]));
    SET9(24, DG(rx2, rypz, +1, rx1, +1, r2ymz, lp2u[4], u[4], mf),
             DG(ry2, rxpz, +1, ry1, +1, r2xmz, lp2u[4], u[4], mf),
             DG(rz2, rxpy, -1, rz1, +1, r2xmy, lp2u[4], u[4], mf),
             t0, t0, t1, t1, t2, t2);
    ox = 2; oz = 1; mf = MS(wm3o12);  /* a8p ccp */
    t0 = A_(N_(M_(rz2, A_(u[8], u[4]))), mf); t1 = A_(N_(M_(rz2, A_(lp2u[8], lp2u[4]))), mf);
    t2 = M_(rxy, A_(lpu[8], lpu[4])); t3 = N_(M_(rxz, S_(lmu[8], lmu[4]))); t4 = N_(M_(ryz, S_(lmu[8], lmu[4])));
    SET9(25, t0, t0, t1, t2, t2, t3, N_(t3), t4, N_(t4));
    ox = 2; oz = 2; mf = MS(wm4o8);   /* a9p cfp */
    t0 = M_(rxy, lpu[8]); t1 = M_(rxz, lpu[8]); t2 = M_(ryz, lpu[8]);
    SET9(26, DG(rx2, rypz, +1, rx1, +1, r2ypz, lp2u[8], u[8], mf),
             DG(ry2, rxpz, +1, ry1, +1, r2xpz, lp2u[8], u[8], mf),
             DG(rz2, rxpy, +1, rz1, +1, r2xpy, lp2u[8], u[8], mf),
             t0, t0, t1, t1, t2, t2);
#undef M_
#undef A_
#undef S_
#undef N_
#undef MS
#undef DG
#undef SET9
    return 0;
}

/* Region class of a node, fdfd3d.f:50-115.  Returns 0 interior, 1 PML,
 * 2 Dirichlet, 3 traction-free top (iz = 1, usemin = false, fdfd3d.f:73-85);
 * *izp is the z index the PML form is evaluated at. */
static int FN(region_)(const FN(prm_) *P, int iz, int ix, int iy, int *izp)
{
    int nx = P->nx, ny = P->ny, nz = P->nz, ip = P->ipml;
    int inx = (ix > ip) && (ix <= nx - ip), iny = (iy > ip) && (iy <= ny - ip);
    *izp = iz;
    if (inx && iny && (iz > ip) && (iz <= nz - ip)) return 0;
    if (iz == nz || ix == 1 || ix == nx || iy == 1 || iy == ny) return 2;
    if (iz <= ip) {
        if (iz == 1) {
            if (!P->fst) return 2;
            if (!P->usemin) return 3;
        } else if (!P->fst) {
            return 1;
        }
        if (inx && iny) return 0;
        *izp = ip + 1;
        return 1;
    }
    return 1;
}

/* ASMBLE, asmble.f90:96-601: values of the CSR matrix in the pattern order
 * (node y > x > z, then row component, then neighbour k ascending, then
 * column component).  Returns 0, or the reference's error conditions. */
int FN(ora_assemble_)(int nx, int ny, int nz, R dx, R om_re, R om_im,
                      int fst, int usemin, R pmld, R pmlf, int ipml,
                      const R *lam, const R *mu, const R *rho, R *a)
{
    FN(prm_) P;
    P.nx = nx; P.ny = ny; P.nz = nz; P.ipml = ipml; P.fst = fst; P.usemin = usemin;
    P.dx = dx; P.pmld = pmld; P.pmlf = pmlf; P.omega = FN(mk_)(om_re, om_im);
    const CT *L = (const CT *)lam, *M = (const CT *)mu;
    CT *out = (CT *)a;
    long pos = 0;
    CT blk[27][9];
    for (int iy = 1; iy <= ny; ++iy)
        for (int ix = 1; ix <= nx; ++ix)
            for (int iz = 1; iz <= nz; ++iz) {
                int izp, mode = FN(region_)(&P, iz, ix, iy, &izp);
                if (FN(node_blocks_)(&P, L, M, rho, iz, ix, iy, mode, izp, blk)) return 1;
                int has[27];
                for (int oy = -1; oy <= 1; ++oy)
                    for (int ox = -1; ox <= 1; ++ox)
                        for (int oz = -1; oz <= 1; ++oz)
                            has[9 * (oy + 1) + 3 * (ox + 1) + (oz + 1)] =
                                (iy + oy >= 1 && iy + oy <= ny && ix + ox >= 1 && ix + ox <= nx &&
                                 iz + oz >= 1 && iz + oz <= nz);
                for (int r = 0; r < 3; ++r)
                    for (int k = 0; k < 27; ++k) {
                        if (!has[k]) continue;
                        for (int c = 0; c < 3; ++c) out[pos++] = blk[k][3 * r + c];
                    }
            }
    return 0;
}

#undef CT
#undef FN
