// Geometric nested dissection + multifrontal symbolic analysis (host).
#include "wt_symbolic.h"

#include <algorithm>
#include <cstdio>
#include <thread>

#include "wt_common.h"

namespace {

struct Box {
    int x0, x1, y0, y1, z0, z1;
    int64_t vol() const { return (int64_t)(x1 - x0) * (y1 - y0) * (z1 - z0); }
};

struct Builder {
    wt_symbolic& S;
    int leaf;
    int64_t next = 0;  // next free position in the elimination order

    // Assigns elimination positions to the nodes of box b (y > x > z order,
    // three components per node) starting at `next`.
    void number(const Box& b)
    {
        const int nx = S.nx, nz = S.nz;
        for (int y = b.y0; y < b.y1; ++y)
            for (int x = b.x0; x < b.x1; ++x)
                for (int z = b.z0; z < b.z1; ++z) {
                    int64_t g = z + (int64_t)x * nz + (int64_t)y * nx * nz;
                    for (int c = 0; c < 3; ++c) {
                        S.perm[3 * g + c] = next;
                        S.iperm[next] = 3 * g + c;
                        ++next;
                    }
                }
    }

    int build(const Box& b, int depth)
    {
        int dims[3] = {b.x1 - b.x0, b.y1 - b.y0, b.z1 - b.z0};
        // longest axis; ties prefer y, then x, then z
        int ax = 1;
        if (dims[0] > dims[ax]) ax = 0;
        if (dims[2] > dims[ax]) ax = 2;
        int kids[2] = {-1, -1};
        Box sep = b;
        if (b.vol() > leaf && dims[ax] >= 3) {
            int lo = ax == 0 ? b.x0 : ax == 1 ? b.y0 : b.z0;
            int s = lo + dims[ax] / 2;
            Box c0 = b, c1 = b;
            if (ax == 0) { c0.x1 = s; c1.x0 = s + 1; sep.x0 = s; sep.x1 = s + 1; }
            if (ax == 1) { c0.y1 = s; c1.y0 = s + 1; sep.y0 = s; sep.y1 = s + 1; }
            if (ax == 2) { c0.z1 = s; c1.z0 = s + 1; sep.z0 = s; sep.z1 = s + 1; }
            kids[0] = build(c0, depth + 1);
            kids[1] = build(c1, depth + 1);
        }
        wt_front f{};
        f.level = depth;
        f.parent = -1;
        f.child[0] = kids[0];
        f.child[1] = kids[1];
        f.slot = 0;
        f.pstart = next;
        f.p = 3 * sep.vol();
        f.box[0] = b.x0; f.box[1] = b.x1; f.box[2] = b.y0;
        f.box[3] = b.y1; f.box[4] = b.z0; f.box[5] = b.z1;
        number(sep);
        int id = (int)S.fronts.size();
        S.fronts.push_back(f);
        for (int k = 0; k < 2; ++k)
            if (kids[k] >= 0) {
                S.fronts[kids[k]].parent = id;
                S.fronts[kids[k]].slot = k;
            }
        return id;
    }
};

}  // namespace

int wt_symbolic_build(int nx, int ny, int nz, int leaf_nodes, int nb, const int32_t* rowptr,
                      const int32_t* colidx, wt_symbolic& S)
{
    if (nx < 1 || ny < 1 || nz < 1 || nb < 1 || leaf_nodes < 1) return WT_ERR_ARG;
    S = wt_symbolic();
    S.nx = nx; S.ny = ny; S.nz = nz; S.nb = nb;
    S.n = 3LL * nx * ny * nz;
    S.perm.assign(S.n, -1);
    S.iperm.assign(S.n, -1);
    Builder B{S, leaf_nodes};
    B.build(Box{0, nx, 0, ny, 0, nz}, 0);
    if (B.next != S.n) return WT_ERR_STATE;

    // update sets: the halo of each front's box, ordered by elimination
    // position (the three components of a node are consecutive)
    const int nf = (int)S.fronts.size();
    int64_t uo = 0;
    std::vector<int64_t> halo;
    for (int f = 0; f < nf; ++f) {
        wt_front& F = S.fronts[f];
        halo.clear();
        const int* b = F.box;
        for (int y = std::max(b[2] - 1, 0); y < std::min(b[3] + 1, ny); ++y)
            for (int x = std::max(b[0] - 1, 0); x < std::min(b[1] + 1, nx); ++x)
                for (int z = std::max(b[4] - 1, 0); z < std::min(b[5] + 1, nz); ++z) {
                    if (x >= b[0] && x < b[1] && y >= b[2] && y < b[3] && z >= b[4] && z < b[5]) continue;
                    int64_t g = z + (int64_t)x * nz + (int64_t)y * nx * nz;
                    halo.push_back(S.perm[3 * g]);
                }
        std::sort(halo.begin(), halo.end());
        F.u = 3 * (int64_t)halo.size();
        F.uoff = uo;
        for (int64_t h : halo)
            for (int c = 0; c < 3; ++c) S.upd_idx.push_back(h + c);
        uo += F.u;
        if (!halo.empty() && halo.front() < F.pstart + F.p) return WT_ERR_STATE;
    }

    // child -> parent positions
    S.relmap.assign(uo, -1);
    for (int f = 0; f < nf; ++f) {
        wt_front& F = S.fronts[f];
        F.moff = F.uoff;
        if (F.parent < 0) {
            if (F.u != 0) return WT_ERR_STATE;
            continue;
        }
        const wt_front& P = S.fronts[F.parent];
        const int64_t* pu = S.upd_idx.data() + P.uoff;
        for (int64_t i = 0; i < F.u; ++i) {
            int64_t v = S.upd_idx[F.uoff + i];
            int64_t pos;
            if (v >= P.pstart && v < P.pstart + P.p) {
                pos = v - P.pstart;
            } else {
                const int64_t* it = std::lower_bound(pu, pu + P.u, v);
                if (it == pu + P.u || *it != v) return WT_ERR_STATE;
                pos = P.p + (it - pu);
            }
            S.relmap[F.moff + i] = (int32_t)pos;
        }
    }

    // levels and arena offsets
    int maxlev = 0;
    for (auto& F : S.fronts) maxlev = std::max(maxlev, F.level);
    S.levels.assign(maxlev + 1, {});
    for (int f = 0; f < nf; ++f) S.levels[S.fronts[f].level].push_back(f);
    int64_t fo = 0, to = 0, po = 0;
    for (int f = 0; f < nf; ++f) {
        wt_front& F = S.fronts[f];
        F.Loff = fo;
        fo += (F.p + F.u) * F.p;
        F.Uoff = fo;
        fo += F.p * F.u;
        int64_t nt = (F.p + nb - 1) / nb;
        F.tileoff = to;
        to += 2 * nt * nb * nb;
        F.pivoff = po;
        po += F.p;
        S.max_p = std::max(S.max_p, F.p);
        S.max_u = std::max(S.max_u, F.u);
        double p = (double)F.p, u = (double)F.u;
        S.flops += 8.0 * (p * p * p / 3.0 + p * p * u + p * u * u);
    }
    S.factor_len = fo;
    S.tile_len = to;
    S.piv_len = po;
    for (int l = 0; l <= maxlev; ++l) {
        int64_t co = 0, wo = 0;
        for (int f : S.levels[l]) {
            S.fronts[f].CBoff = co;
            co += S.fronts[f].u * S.fronts[f].u;
            S.fronts[f].Woff = wo;
            wo += S.fronts[f].u;
        }
        S.cb_len[l & 1] = std::max(S.cb_len[l & 1], co);
        S.w_len[l & 1] = std::max(S.w_len[l & 1], wo);
    }

    // destination of every CSR entry in the factor arena: the entry (i, j)
    // belongs to the front that eliminates min(perm i, perm j)
    if (rowptr && colidx) {
        S.nnz = rowptr[S.n];
        S.dst.assign(S.nnz, -1);
        unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<int> bad(nth, 0);
        auto work = [&](unsigned t) {
            for (int f = (int)t; f < nf; f += (int)nth) {
                const wt_front& F = S.fronts[f];
                const int64_t* uf = S.upd_idx.data() + F.uoff;
                const int64_t pend = F.pstart + F.p, ld = F.p + F.u;
                auto local = [&](int64_t v) -> int64_t {
                    if (v >= F.pstart && v < pend) return v - F.pstart;
                    const int64_t* it = std::lower_bound(uf, uf + F.u, v);
                    if (it == uf + F.u || *it != v) return -1;
                    return F.p + (it - uf);
                };
                auto where = [&](int64_t r, int64_t c) -> int64_t {
                    return c < F.p ? F.Loff + c * ld + r : F.Uoff + (c - F.p) * F.p + r;
                };
                for (int64_t ip = F.pstart; ip < pend; ++ip) {
                    int64_t io = S.iperm[ip];
                    int64_t r = ip - F.pstart;
                    for (int64_t e = rowptr[io]; e < rowptr[io + 1]; ++e) {
                        int64_t jo = colidx[e], jp = S.perm[jo];
                        if (jp < F.pstart) continue;  // a descendant owns it
                        int64_t c = local(jp);
                        if (c < 0) { bad[t] = 1; continue; }
                        S.dst[e] = where(r, c);
                        if (jp >= pend) {
                            // transposed entry (jo, io) sits in row jo
                            const int32_t* cb = colidx + rowptr[jo];
                            const int32_t* ce = colidx + rowptr[jo + 1];
                            const int32_t* it = std::lower_bound(cb, ce, (int32_t)io);
                            if (it == ce || *it != io) { bad[t] = 1; continue; }
                            S.dst[rowptr[jo] + (it - cb)] = where(c, r);
                        }
                    }
                }
            }
        };
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nth; ++t) th.emplace_back(work, t);
        for (auto& x : th) x.join();
        for (int b : bad)
            if (b) return WT_ERR_PATTERN;
        for (int64_t e = 0; e < S.nnz; ++e)
            if (S.dst[e] < 0) return WT_ERR_PATTERN;
    }
    return WT_OK;
}

int wt_pattern_host(int nx, int ny, int nz, int32_t row_base, int32_t col_base, int32_t* rowptr, int32_t* colidx)
{
    if (nx < 1 || ny < 1 || nz < 1 || !rowptr || !colidx) return WT_ERR_ARG;
    int64_t e = 0, row = 0;
    for (int iy = 0; iy < ny; ++iy)
        for (int ix = 0; ix < nx; ++ix)
            for (int iz = 0; iz < nz; ++iz) {
                const int64_t g = iz + (int64_t)ix * nz + (int64_t)iy * nx * nz;
                for (int r = 0; r < 3; ++r) {
                    rowptr[row++] = (int32_t)(e + row_base);
                    for (int oy = -1; oy <= 1; ++oy) {
                        if (iy + oy < 0 || iy + oy >= ny) continue;
                        for (int ox = -1; ox <= 1; ++ox) {
                            if (ix + ox < 0 || ix + ox >= nx) continue;
                            for (int oz = -1; oz <= 1; ++oz) {
                                if (iz + oz < 0 || iz + oz >= nz) continue;
                                const int64_t gn = g + oz + (int64_t)ox * nz + (int64_t)oy * nx * nz;
                                for (int c = 0; c < 3; ++c) colidx[e++] = (int32_t)(3 * gn + c + col_base);
                            }
                        }
                    }
                }
            }
    rowptr[row] = (int32_t)(e + row_base);
    return WT_OK;
}

int wt_symbolic_union(const std::vector<const wt_symbolic*>& parts, wt_symbolic& U)
{
    U = wt_symbolic();
    if (parts.empty()) return WT_ERR_ARG;
    U.nb = parts[0]->nb;
    int maxlev = 0;
    int64_t no = 0, nnzo = 0, fo = 0, to = 0, po = 0, uo = 0, mo = 0;
    int32_t foff = 0;
    size_t nf = 0, nu = 0, nm = 0;
    for (const wt_symbolic* P : parts) {
        if (P->nb != U.nb || (int64_t)P->dst.size() != P->nnz) return WT_ERR_ARG;
        nf += P->fronts.size();
        nu += P->upd_idx.size();
        nm += P->relmap.size();
        nnzo += P->nnz;
        no += P->n;
    }
    U.fronts.reserve(nf);
    U.upd_idx.reserve(nu);
    U.relmap.reserve(nm);
    U.dst.reserve((size_t)nnzo);
    U.perm.reserve((size_t)no);
    U.iperm.reserve((size_t)no);
    no = nnzo = 0;
    for (const wt_symbolic* P : parts) {
        for (const wt_front& F0 : P->fronts) {
            wt_front F = F0;
            if (F.parent >= 0) F.parent += foff;
            for (int k = 0; k < 2; ++k)
                if (F.child[k] >= 0) F.child[k] += foff;
            F.pstart += no;
            F.uoff += uo;
            F.moff += mo;
            F.Loff += fo;
            F.Uoff += fo;
            F.tileoff += to;
            F.pivoff += po;
            maxlev = std::max(maxlev, F.level);
            U.fronts.push_back(F);
        }
        for (int64_t v : P->perm) U.perm.push_back(v + no);
        for (int64_t v : P->iperm) U.iperm.push_back(v + no);
        for (int64_t v : P->upd_idx) U.upd_idx.push_back(v + no);
        U.relmap.insert(U.relmap.end(), P->relmap.begin(), P->relmap.end());
        for (int64_t v : P->dst) U.dst.push_back(v + fo);
        no += P->n;
        nnzo += P->nnz;
        fo += P->factor_len;
        to += P->tile_len;
        po += P->piv_len;
        uo += (int64_t)P->upd_idx.size();
        mo += (int64_t)P->relmap.size();
        foff += (int32_t)P->fronts.size();
        U.max_p = std::max(U.max_p, P->max_p);
        U.max_u = std::max(U.max_u, P->max_u);
        U.flops += P->flops;
    }
    U.n = no;
    U.nnz = nnzo;
    U.factor_len = fo;
    U.tile_len = to;
    U.piv_len = po;
    // levels of the forest, and the per-level contribution-block and
    // solve-update offsets (the same layout rule as one tree)
    U.levels.assign(maxlev + 1, {});
    for (int f = 0; f < (int)U.fronts.size(); ++f) U.levels[U.fronts[f].level].push_back(f);
    for (int l = 0; l <= maxlev; ++l) {
        int64_t co = 0, wo = 0;
        for (int f : U.levels[l]) {
            U.fronts[f].CBoff = co;
            co += U.fronts[f].u * U.fronts[f].u;
            U.fronts[f].Woff = wo;
            wo += U.fronts[f].u;
        }
        U.cb_len[l & 1] = std::max(U.cb_len[l & 1], co);
        U.w_len[l & 1] = std::max(U.w_len[l & 1], wo);
    }
    return WT_OK;
}
