/* Precision-generic body of the Thomson-Haskell restatement, included by
 * hask_oracle.c with
 *   RS      the reference's REAL type (float as compiled, double under
 *           -fdefault-real-8): inputs, single-precision subexpressions,
 *           outputs;
 *   SINRS   the generic SIN of a REAL argument;
 *   FN(x)   function-name suffixing.
 * Every other quantity is DOUBLE PRECISION, as in haskgrn.f.
 *
 * TEST INFRASTRUCTURE ONLY (see hask_oracle.c).
 */

/* Layer-propagator trig/hyperbolic terms for phase z (haskgrn.f:168-187):
 * trig when the horizontal speed exceeds the layer speed, hyperbolic
 * otherwise. */
static void FN(hk_phase)(double cv, double vlay, double z, double *e, double *c, double *s)
{
    if (cv > vlay) {
        *e = 1.0;
        *c = cos(z);
        *s = sin(z);
    } else {
        double x = exp(z);
        *e = -1.0;
        *c = 0.5 * (x + 1.0 / x);
        *s = 0.5 * (x - 1.0 / x);
    }
}

/* b <- A_m b over the first two columns, with the R/I sign pattern of
 * haskgrn.f:231-321 (a-terms of H53 p.21; fluid layer H53 eq 6.3). */
static void FN(hk_layer)(const double *L, int m, double b[8])
{
    /* L holds per layer: r1 r2 aag aag1 tt xh xh1 vsdisc e1 c01 s01 e2 c02 s02 */
    const double *q = L + 14 * m;
    const double r1 = q[0], r2 = q[1], ag = q[2], ag1 = q[3], tm = q[4], x = q[5], y = q[6], vsd = q[7];
    const double e1 = q[8], cp = q[9], s01 = q[10], e2 = q[11], cq = q[12], s02 = q[13];
    const double pd = s01 / r1, qd = s02 / r2, pm = e1 * s01 * r1, qm = e2 * s02 * r2;
    double a11, a12, a13, a14, a21, a22, a23, a24, a31, a32, a33, a34, a41, a42, a43, a44;
    if (vsd != 0.0) {
        a11 = ag * cp - ag1 * cq;
        a12 = ag1 * pd + ag * qm;
        a13 = (cq - cp) / tm;
        a14 = (pd + qm) / tm;
        a21 = -ag * pm - ag1 * qd;
        a22 = ag * cq - ag1 * cp;
        a23 = (pm + qd) / tm;
        a24 = a13;
        a31 = tm * ag * ag1 * (cp - cq);
        a32 = tm * (y * pd + x * qm);
        a33 = a22;
        a34 = a12;
        a41 = tm * (x * pm + y * qd);
        a42 = a31;
        a43 = a21;
        a44 = a11;
    } else {
        a11 = 0.0; a12 = -pd; a13 = -cp / tm; a14 = 0.0;
        a21 = 0.0; a22 = cp; a23 = pm / tm; a24 = 0.0;
        a31 = 0.0; a32 = tm * pd; a33 = cp; a34 = 0.0;
        a41 = 0.0; a42 = 0.0; a43 = 0.0; a44 = 0.0;
    }
    const double b11 = b[0], b12 = b[1], b21 = b[2], b22 = b[3], b31 = b[4], b32 = b[5], b41 = b[6], b42 = b[7];
    b[0] = a11 * b11 - a12 * b21 + a13 * b31 - a14 * b41;
    b[1] = a11 * b12 + a12 * b22 + a13 * b32 + a14 * b42;
    b[2] = a21 * b11 + a22 * b21 + a23 * b31 + a24 * b41;
    b[3] = -a21 * b12 + a22 * b22 - a23 * b32 + a24 * b42;
    b[4] = a31 * b11 - a32 * b21 + a33 * b31 - a34 * b41;
    b[5] = a31 * b12 + a32 * b22 + a33 * b32 + a34 * b42;
    b[6] = a41 * b11 + a42 * b21 + a43 * b31 + a44 * b41;
    b[7] = -a41 * b12 + a42 * b22 - a43 * b32 + a44 * b42;
}

/* HASKGRN (haskgrn.f:31-672) for one source.
 *   nl layers (0-based arrays; layer nl-1 is the half-space), top depths h1d,
 *   vp1d/vs1d the speeds that set the horizontal speed, vp1disc/vs1disc the
 *   speeds of the layer terms, rh1d densities; ain incidence angle
 *   (radians), is_s = 1 for an incident S wave ('ps'), df = omega; nz depths
 *   dept(iz) = dz (iz-1) + zorig.
 * Out: ugrn/wgrn interleaved complex [2*nz]; *pic = omega / cv.
 * Returns 0, 1 when a depth lies above h1d(1) (haskgrn.f:437-443; outputs
 * from that depth on are left as they were), 2 for nl < 2 with nz > 1 (the
 * reference would read h1d(0)). */
int FN(ora_haskgrn)(int nl, const RS *vp1d, const RS *vs1d, const RS *rh1d, const RS *h1d,
                    const RS *vp1disc, const RS *vs1disc, RS ain, int is_s, RS df, int nz, RS dz,
                    RS zorig, RS *ugrn, RS *wgrn, double *pic_out)
{
    if (nl < 1 || nz < 1) return 2;
    if (nl < 2 && nz > 1) return 2;
    double *L = (double *)calloc((size_t)14 * nl, sizeof(double));
    double *h = (double *)calloc((size_t)nl, sizeof(double));
    if (!L || !h) { free(L); free(h); return 3; }
    const int N = nl - 1; /* 0-based index of the half-space */
    /* haskgrn.f:104-116: horizontal speed from the continuous speeds, in REAL */
    const double cv = (double)(RS)((is_s ? vs1d[N] : vp1d[N]) / SINRS(ain));
    const double pic = (double)df / cv;
    if (pic_out) *pic_out = pic;
    for (int i = 0; i < N; ++i) h[i] = (double)(RS)(h1d[i + 1] - h1d[i]);
    /* haskgrn.f:124-189 */
    for (int m = 0; m < nl; ++m) {
        double *q = L + 14 * m;
        const double cova = cv / (double)vp1disc[m];
        const double covb = vs1disc[m] != (RS)0 ? cv / (double)vs1disc[m] : 1.0;
        double x = cova * cova - 1.0;
        q[0] = copysign(sqrt(fabs(x)), x);
        x = covb * covb - 1.0;
        q[1] = copysign(sqrt(fabs(x)), x);
        q[2] = 2.0 / (covb * covb);
        q[3] = q[2] - 1.0;
        q[7] = (double)vs1disc[m];
        if (m < N) {
            q[4] = (double)rh1d[m] * cv * cv;
            q[5] = q[2] * q[2];
            q[6] = q[3] * q[3];
            const double hc = pic * h[m];
            FN(hk_phase)(cv, (double)vp1disc[m], hc * q[0], &q[8], &q[9], &q[10]);
            FN(hk_phase)(cv, (double)vs1disc[m], hc * q[1], &q[11], &q[12], &q[13]);
        }
    }
    /* scaled inverse E of the half-space (haskgrn.f:200-209), REAL parts in REAL */
    const double *qN = L + 14 * N;
    const double xe = -qN[3] * cv * cv;
    const double e11 = (double)((RS)2 * vs1disc[N] * vs1disc[N]);
    const double e13 = (double)((RS)-1 / rh1d[N]);
    const double e22 = xe / qN[0], e24 = e13 / qN[0], e31 = xe / qN[1], e33 = -e13 / qN[1];
    const double e42 = -e11, e44 = e13;
    double b[8] = {1.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0};
    for (int m = 0; m < N; ++m) FN(hk_layer)(L, m, b);
    double c11 = e11 * b[0] + e13 * b[4], c12 = e11 * b[1] + e13 * b[5];
    double c21 = e22 * b[2] + e24 * b[6], c22 = e22 * b[3] + e24 * b[7];
    double c31 = e31 * b[0] + e33 * b[4], c32 = e31 * b[1] + e33 * b[5];
    double c41 = e42 * b[2] + e44 * b[6], c42 = e42 * b[3] + e44 * b[7];
    const double scl1 = (double)((RS)-1 / (vp1disc[N] * vp1disc[N]));
    const double scl2 = (double)((RS)-1 / ((RS)2 * vs1disc[N] * vs1disc[N]));
    c11 *= scl1; c12 *= scl1; c21 *= scl1; c22 *= scl1;
    c31 *= scl2; c32 *= scl2; c41 *= scl2; c42 *= scl2;
    const double x = -c11 * c42 + c21 * c32 + c31 * c22 - c41 * c12;
    const double y = c11 * c32 + c21 * c42 - c31 * c12 - c41 * c22;
    const double ag = x * x + y * y;
    double wr, wi, ur, ui;
    if (!is_s) {
        const double tm = 2.0 * cv / ((double)vp1disc[N] * ag);
        wr = tm * (c31 * x - c41 * y);
        wi = tm * (-c31 * y - c41 * x);
        ur = tm * (c42 * x - c32 * y);
        ui = tm * (-c42 * y - c32 * x);
    } else {
        const double tm = cv / ((double)vs1disc[N] * ag);
        wr = tm * (-c11 * x + c21 * y);
        wi = tm * (c11 * y + c21 * x);
        ur = tm * (-c22 * x + c12 * y);
        ui = tm * (c22 * y + c12 * x);
    }
    const double qrw = wr, qiw = wi, qru = ur, qiu = ui;
    ugrn[0] = (RS)qru; ugrn[1] = (RS)qiu;
    wgrn[0] = (RS)qrw; wgrn[1] = (RS)qiw;
    /* haskgrn.f:425-655: walk down the grid; layer terms are overwritten as
     * the walk goes, exactly as the reference does (1-based layer numbers
     * below, like the reference's) */
    int ibovo = 1, ifin = 0;
    int mstale = nl; /* the DO variable m after haskgrn.f:231 */
    int rc = 0;
    for (int iz = 2; iz <= nz; ++iz) {
        const double dept = (double)(dz * (RS)(iz - 1) + zorig);
        int i;
        for (i = 2; i <= nl; ++i)
            if (dept < (double)h1d[i - 1]) break;
        if (i > nl) i = nl;
        const int ibov = i - 1;
        const double dloc = dept - (double)h1d[i - 2];
        if (dloc < 0.0) { rc = 1; break; }
        if (ibov != ibovo) {
            ibovo = ibov;
            if (ibov > 1) {
                ifin += 1;
                double *q = L + 14 * (ifin - 1);
                const double hc = pic * h[ifin - 1];
                /* the branch tests read the stale DO variable (haskgrn.f:454,466) */
                FN(hk_phase)(cv, (double)vp1disc[mstale - 1], hc * q[0], &q[8], &q[9], &q[10]);
                FN(hk_phase)(cv, (double)vs1disc[mstale - 1], hc * q[1], &q[11], &q[12], &q[13]);
            }
        }
        {
            double *q = L + 14 * (ibov - 1);
            const double hc = pic * dloc;
            FN(hk_phase)(cv, (double)vp1disc[ibov - 1], hc * q[0], &q[8], &q[9], &q[10]);
            FN(hk_phase)(cv, (double)vs1disc[ibov - 1], hc * q[1], &q[11], &q[12], &q[13]);
        }
        double bb[8] = {1.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0};
        for (int m = 0; m < ibov; ++m) FN(hk_layer)(L, m, bb);
        mstale = ibov + 1;
        const double u_r = bb[0] * qru - bb[1] * qiw;
        const double u_i = bb[0] * qiu + bb[1] * qrw;
        const double w_r = -bb[2] * qiu + bb[3] * qrw;
        const double w_i = bb[2] * qru + bb[3] * qiw;
        ugrn[2 * (iz - 1)] = (RS)u_r; ugrn[2 * (iz - 1) + 1] = (RS)u_i;
        wgrn[2 * (iz - 1)] = (RS)w_r; wgrn[2 * (iz - 1) + 1] = (RS)w_i;
    }
    free(L);
    free(h);
    return rc;
}
