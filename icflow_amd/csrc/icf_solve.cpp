// icf_solve.cpp -- SolveDouble3: the linear solve of one Newton iteration on the host, behind
// the signature the reference's C# LinearSystem.Solve calls (LinearSystem.cs:45-64;
// PardisoLoader2/dllmain.cpp:29-79).  It is NOT MKL PARDISO, which this environment does not
// have: it is a sparse LDL^T of the symmetric matrix given by its upper-triangular 3x3 BCSR
// (PARDISO's mtype -2 input with iparm[34] = 1, zero-based, and iparm[36] = 3, block size 3).
//
//   ordering   of the block graph, whichever of two gives the smaller factor: nested
//              dissection (each subgraph is split by the middle level set of a breadth-first
//              search from a pseudo-peripheral vertex -- the level set is a separator; vertices of
//              it without a neighbour in the next level move to the near side -- both sides are
//              ordered first and the separator last; subgraphs of at most 64 blocks keep their
//              search order), which suits compact 3D meshes, or reverse Cuthill-McKee, which
//              suits long thin ones (beams);
//   symbolic   elimination tree and column counts of L (block level);
//   numeric    up-looking block LDL^T: row k of L from a sparse forward solve along the
//              elimination tree, D_k a dense symmetric 3x3 block.  A pivot block that is
//              singular to working precision gets 1e-8 of the matrix's largest entry added to
//              its diagonal (PARDISO perturbs small pivots by 10^-iparm[9] with iparm[9] = 8 in
//              the reference's settings), and the solution then takes up to two steps of
//              iterative refinement;
//   solve      L y = P b, D z = y, L^T w = z, x = P^T w.
//
// The ordering and symbolic analysis of the last sparsity pattern are kept (the reference
// re-analyses on every call; the result does not depend on this).
#include "../../include/icflow_host.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

namespace {

// 3x3 blocks, row-major
inline void mm(const double *a, const double *b, double *c)   // c = a b
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) c[3 * i + j] = a[3 * i] * b[j] + a[3 * i + 1] * b[3 + j] + a[3 * i + 2] * b[6 + j];
}
inline void mtm(const double *a, const double *b, double *c)  // c = a^T b
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) c[3 * i + j] = a[i] * b[j] + a[3 + i] * b[3 + j] + a[6 + i] * b[6 + j];
}
inline void sub_mm(double *c, const double *a, const double *b)   // c -= a b
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) c[3 * i + j] -= a[3 * i] * b[j] + a[3 * i + 1] * b[3 + j] + a[3 * i + 2] * b[6 + j];
}
inline void transpose(const double *a, double *t)
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) t[3 * j + i] = a[3 * i + j];
}

// Inverse of a symmetric 3x3 block (symmetrised first); false if singular to working precision.
bool inv3(const double *d, double *o, double tiny)
{
    const double a = d[0], b = 0.5 * (d[1] + d[3]), c = 0.5 * (d[2] + d[6]);
    const double e = d[4], f = 0.5 * (d[5] + d[7]), g = d[8];
    const double c00 = e * g - f * f, c01 = c * f - b * g, c02 = b * f - c * e;
    const double det = a * c00 + b * c01 + c * c02;
    const double scale = std::max(std::max(std::fabs(a), std::fabs(e)), std::fabs(g));
    if (!(std::fabs(det) > tiny * scale * scale) || !std::isfinite(det)) return false;
    const double id = 1.0 / det;
    o[0] = c00 * id;
    o[1] = o[3] = c01 * id;
    o[2] = o[6] = c02 * id;
    o[4] = (a * g - c * c) * id;
    o[5] = o[7] = (b * c - a * f) * id;
    o[8] = (a * e - b * b) * id;
    return true;
}

// ---- nested dissection of the block graph ------------------------------------------------------

struct Graph {
    std::vector<int> p, j;   // symmetric adjacency, no self loops
};

struct ND {
    const Graph &g;
    std::vector<int> mark, level, queue;
    int next_label = 1;
    std::vector<int> out;
    explicit ND(const Graph &gr, int n) : g(gr), mark(n, 0), level(n, -1), queue(n) {}

    // breadth-first levels within label; returns the number reached, last vertex in `far`
    int bfs(int s, int label, int *far, int *depth)
    {
        int qh = 0, qt = 0;
        queue[qt++] = s;
        level[s] = 0;
        int last = s;
        while (qh < qt) {
            const int v = queue[qh++];
            last = v;
            for (int q = g.p[v]; q < g.p[v + 1]; q++) {
                const int w = g.j[q];
                if (mark[w] == label && level[w] < 0) {
                    level[w] = level[v] + 1;
                    queue[qt++] = w;
                }
            }
        }
        *far = last;
        *depth = level[last];
        return qt;
    }
    void reset_levels(int count)
    {
        for (int i = 0; i < count; i++) level[queue[i]] = -1;
    }

    void order(std::vector<int> &V)
    {
        const int n = (int)V.size();
        if (n == 0) return;
        if (n <= 64) {
            out.insert(out.end(), V.begin(), V.end());
            return;
        }
        const int label = next_label++;
        for (int v : V) mark[v] = label;
        // pseudo-peripheral start: two sweeps
        int far = V[0], depth = 0;
        int reached = bfs(V[0], label, &far, &depth);
        reset_levels(reached);
        int s = far;
        reached = bfs(s, label, &far, &depth);
        if (reached < n) {
            // several components: each is ordered on its own, one after the other (no
            // separator between them)
            reset_levels(reached);
            std::vector<std::vector<int>> comps;
            for (int v : V) {
                if (mark[v] != label) continue;
                int f, d;
                const int c = bfs(v, label, &f, &d);
                comps.emplace_back(queue.begin(), queue.begin() + c);
                reset_levels(c);
                for (int w : comps.back()) mark[w] = 0;
            }
            std::vector<int>().swap(V);
            for (auto &c : comps) order(c);
            return;
        }
        if (depth < 2) {                  // too small a diameter to split
            reset_levels(reached);
            for (int v : V) mark[v] = 0;
            out.insert(out.end(), V.begin(), V.end());
            return;
        }
        // the level where the cumulative count reaches half, kept inside (0, depth)
        std::vector<int> cnt(depth + 1, 0);
        for (int i = 0; i < reached; i++) cnt[level[queue[i]]]++;
        int m = 1, acc = cnt[0];
        while (m < depth - 1 && acc + cnt[m] < n / 2) acc += cnt[m++];
        std::vector<int> A, B, S;
        A.reserve(acc);
        for (int i = 0; i < reached; i++) {
            const int v = queue[i], l = level[v];
            if (l < m) {
                A.push_back(v);
            } else if (l > m) {
                B.push_back(v);
            } else {
                bool touches = false;      // a neighbour on the far side?
                for (int q = g.p[v]; q < g.p[v + 1] && !touches; q++) {
                    const int w = g.j[q];
                    touches = mark[w] == label && level[w] == m + 1;
                }
                (touches ? S : A).push_back(v);
            }
        }
        reset_levels(reached);
        for (int v : V) mark[v] = 0;
        std::vector<int>().swap(V);
        order(A);
        order(B);
        out.insert(out.end(), S.begin(), S.end());
    }
};

// Reverse Cuthill-McKee: breadth-first from a pseudo-peripheral vertex of each component,
// neighbours in increasing degree, the whole order reversed.
std::vector<int> rcm(const Graph &g, int n)
{
    std::vector<int> order;
    order.reserve(n);
    std::vector<int> level(n, -1), queue(n);
    std::vector<char> done(n, 0);
    auto deg = [&](int v) { return g.p[v + 1] - g.p[v]; };
    auto bfs_far = [&](int s) {   // the last vertex reached from s (levels reset afterwards)
        int qh = 0, qt = 0, last = s;
        queue[qt++] = s;
        level[s] = 0;
        while (qh < qt) {
            const int v = queue[qh++];
            last = v;
            for (int q = g.p[v]; q < g.p[v + 1]; q++)
                if (!done[g.j[q]] && level[g.j[q]] < 0) {
                    level[g.j[q]] = level[v] + 1;
                    queue[qt++] = g.j[q];
                }
        }
        for (int i = 0; i < qt; i++) level[queue[i]] = -1;
        return last;
    };
    std::vector<int> nb;
    for (int s0 = 0; s0 < n; s0++) {
        if (done[s0]) continue;
        const int s = bfs_far(bfs_far(s0));
        const size_t first = order.size();
        order.push_back(s);
        done[s] = 1;
        for (size_t h = first; h < order.size(); h++) {
            const int v = order[h];
            nb.clear();
            for (int q = g.p[v]; q < g.p[v + 1]; q++)
                if (!done[g.j[q]]) {
                    done[g.j[q]] = 1;
                    nb.push_back(g.j[q]);
                }
            std::sort(nb.begin(), nb.end(), [&](int a, int b) { return deg(a) < deg(b) || (deg(a) == deg(b) && a < b); });
            order.insert(order.end(), nb.begin(), nb.end());
        }
    }
    std::reverse(order.begin(), order.end());
    return order;
}

// Number of off-diagonal blocks of L for ordering perm (new -> old): elimination tree and
// column counts of the permuted upper pattern (the row-subtree walk).
long long factor_blocks(const Graph &g, int n, const std::vector<int> &perm)
{
    std::vector<int> iperm(n), parent(n, -1), flag(n, -1);
    for (int k = 0; k < n; k++) iperm[perm[k]] = k;
    long long total = 0;
    for (int k = 0; k < n; k++) {
        flag[k] = k;
        const int v = perm[k];
        for (int q = g.p[v]; q < g.p[v + 1]; q++) {
            int i = iperm[g.j[q]];
            if (i >= k) continue;
            for (; flag[i] != k; i = parent[i]) {
                if (parent[i] == -1) parent[i] = k;
                total++;
                flag[i] = k;
            }
        }
    }
    return total;
}

// ---- analysis (kept for the last pattern) ------------------------------------------------------

struct Analysis {
    int n = -1;
    std::vector<int> ia, ja;       // the pattern analysed
    std::vector<int> perm, iperm;  // new -> old, old -> new
    // the permuted upper matrix by column: rows i < k of column k, the source block and
    // whether it is transposed; the diagonal's source per column
    std::vector<int> cp, ci, csrc;
    std::vector<char> ctr;
    std::vector<int> dsrc;
    std::vector<int> parent;       // elimination tree
    std::vector<long long> lp;     // column starts of L
    long long lnz = 0;
};

std::mutex g_mu;
Analysis g_an;

int analyse(Analysis &an, const int *ia, const int *ja, int n)
{
    an = Analysis();
    an.n = n;
    an.ia.assign(ia, ia + n + 1);
    an.ja.assign(ja, ja + ia[n]);
    // block graph (symmetric, no diagonal)
    Graph g;
    g.p.assign(n + 1, 0);
    for (int r = 0; r < n; r++)
        for (int q = ia[r]; q < ia[r + 1]; q++)
            if (ja[q] != r) {
                g.p[r + 1]++;
                g.p[ja[q] + 1]++;
            }
    for (int r = 0; r < n; r++) g.p[r + 1] += g.p[r];
    g.j.resize(g.p[n]);
    {
        std::vector<int> at(g.p.begin(), g.p.end() - 1);
        for (int r = 0; r < n; r++)
            for (int q = ia[r]; q < ia[r + 1]; q++)
                if (ja[q] != r) {
                    g.j[at[r]++] = ja[q];
                    g.j[at[ja[q]]++] = r;
                }
    }
    ND nd(g, n);
    nd.out.reserve(n);
    std::vector<int> all(n);
    for (int i = 0; i < n; i++) all[i] = i;
    nd.order(all);
    if ((int)nd.out.size() != n) return -3;
    std::vector<int> band = rcm(g, n);
    if ((int)band.size() != n) return -3;
    if (factor_blocks(g, n, band) < factor_blocks(g, n, nd.out)) an.perm.swap(band);
    else an.perm.swap(nd.out);
    an.iperm.assign(n, -1);
    for (int k = 0; k < n; k++) an.iperm[an.perm[k]] = k;
    // permuted upper matrix by column
    an.cp.assign(n + 1, 0);
    an.dsrc.assign(n, -1);
    for (int r = 0; r < n; r++)
        for (int q = ia[r]; q < ia[r + 1]; q++) {
            const int pr = an.iperm[r], pc = an.iperm[ja[q]];
            if (pr == pc) an.dsrc[pr] = q;
            else an.cp[std::max(pr, pc) + 1]++;
        }
    for (int k = 0; k < n; k++) an.cp[k + 1] += an.cp[k];
    an.ci.resize(an.cp[n]);
    an.csrc.resize(an.cp[n]);
    an.ctr.resize(an.cp[n]);
    {
        std::vector<int> at(an.cp.begin(), an.cp.end() - 1);
        for (int r = 0; r < n; r++)
            for (int q = ia[r]; q < ia[r + 1]; q++) {
                const int pr = an.iperm[r], pc = an.iperm[ja[q]];
                if (pr == pc) continue;
                const int col = std::max(pr, pc), row = std::min(pr, pc);
                const int at_ = at[col]++;
                an.ci[at_] = row;
                an.csrc[at_] = q;
                an.ctr[at_] = pr > pc;   // block (r, c) lands below the diagonal: use its transpose
            }
    }
    for (int k = 0; k < n; k++)
        if (an.dsrc[k] < 0) return -1;   // a diagonal block is missing
    // elimination tree and column counts
    an.parent.assign(n, -1);
    std::vector<int> flag(n), cnt(n, 0);
    for (int k = 0; k < n; k++) {
        flag[k] = k;
        for (int q = an.cp[k]; q < an.cp[k + 1]; q++)
            for (int i = an.ci[q]; flag[i] != k; i = an.parent[i]) {
                if (an.parent[i] == -1) an.parent[i] = k;
                cnt[i]++;
                flag[i] = k;
            }
    }
    an.lp.assign(n + 1, 0);
    for (int k = 0; k < n; k++) an.lp[k + 1] = an.lp[k] + cnt[k];
    an.lnz = an.lp[n];
    return 0;
}

bool same_pattern(const Analysis &an, const int *ia, const int *ja, int n)
{
    if (an.n != n || (int)an.ia.size() != n + 1 || an.ia[n] != ia[n]) return false;
    return std::equal(an.ia.begin(), an.ia.end(), ia) && std::equal(an.ja.begin(), an.ja.end(), ja);
}

// y += A x for the symmetric matrix given by its upper 3x3 BCSR
void symv(const int *ia, const int *ja, const double *a, int n, const double *x, double *y)
{
    for (int r = 0; r < n; r++)
        for (int q = ia[r]; q < ia[r + 1]; q++) {
            const int c = ja[q];
            const double *b = a + (size_t)q * 9;
            for (int i = 0; i < 3; i++)
                y[3 * r + i] += b[3 * i] * x[3 * c] + b[3 * i + 1] * x[3 * c + 1] + b[3 * i + 2] * x[3 * c + 2];
            if (c != r)
                for (int j = 0; j < 3; j++)
                    y[3 * c + j] += b[j] * x[3 * r] + b[3 + j] * x[3 * r + 1] + b[6 + j] * x[3 * r + 2];
        }
}

}  // namespace

extern "C" int SolveDouble3(int *ja, int *ia, double *a, const int n, double *b, double *x, int mtype, int iparam4,
                            int dim, int msglvl, int check)
{
    (void)iparam4;   // PARDISO's iterative-direct switch; the reference passes 0
    if (n < 0 || !ia || (n > 0 && (!ja || !a || !b || !x))) return -1;
    if (mtype != -2 && mtype != 2) return -1;   // symmetric matrices only (the reference: -2)
    if (dim != 3) return -1;                     // 3x3 blocks (the reference: 3)
    if (n == 0) return 0;
    if (check) {
        if (ia[0] != 0) return -1;
        for (int r = 0; r < n; r++) {
            if (ia[r + 1] < ia[r]) return -1;
            for (int q = ia[r]; q < ia[r + 1]; q++)
                if (ja[q] < r || ja[q] >= n || (q > ia[r] && ja[q] <= ja[q - 1])) return -1;
        }
    }
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    std::lock_guard<std::mutex> lk(g_mu);
    try {
        Analysis &an = g_an;
        if (!same_pattern(an, ia, ja, n))
            if (int r = analyse(an, ia, ja, n)) {
                an = Analysis();
                return r;
            }
        const auto t1 = clk::now();
        double amax = 0;
        for (long long i = 0; i < (long long)ia[n] * 9; i++) amax = std::max(amax, std::fabs(a[i]));
        const double delta = 1e-8 * (amax > 0 ? amax : 1.0);
        // numeric factorisation
        std::vector<int> li((size_t)an.lnz), lnz(n, 0), flag(n, -1), pattern(n), stack(n);
        if (an.lnz >= (1LL << 31)) return -2;
        std::vector<double> lx((size_t)an.lnz * 9), dinv((size_t)n * 9), y((size_t)n * 9, 0.0);
        int perturbed = 0;
        for (int k = 0; k < n; k++) {
            int top = n;
            flag[k] = k;
            for (int q = an.cp[k]; q < an.cp[k + 1]; q++) {
                int i = an.ci[q];
                const double *src = a + (size_t)an.csrc[q] * 9;
                double *yi = &y[(size_t)i * 9];
                if (an.ctr[q]) {
                    double t[9];
                    transpose(src, t);
                    for (int e = 0; e < 9; e++) yi[e] += t[e];
                } else {
                    for (int e = 0; e < 9; e++) yi[e] += src[e];
                }
                int len = 0;
                for (; flag[i] != k; i = an.parent[i]) {
                    stack[len++] = i;
                    flag[i] = k;
                }
                while (len > 0) pattern[--top] = stack[--len];
            }
            double d[9];
            std::memcpy(d, a + (size_t)an.dsrc[k] * 9, sizeof d);
            for (int t = top; t < n; t++) {
                const int i = pattern[t];
                double zi[9];
                std::memcpy(zi, &y[(size_t)i * 9], sizeof zi);
                std::memset(&y[(size_t)i * 9], 0, sizeof zi);
                const long long p2 = an.lp[i] + lnz[i];
                for (long long p = an.lp[i]; p < p2; p++) sub_mm(&y[(size_t)li[p] * 9], &lx[(size_t)p * 9], zi);
                double lki[9];
                mtm(zi, &dinv[(size_t)i * 9], lki);   // L_ki = Z_i^T D_i^-1
                sub_mm(d, lki, zi);                   // D_k -= L_ki Z_i
                li[p2] = k;
                std::memcpy(&lx[(size_t)p2 * 9], lki, sizeof lki);
                lnz[i]++;
            }
            if (!inv3(d, &dinv[(size_t)k * 9], 1e-13)) {
                d[0] += d[0] >= 0 ? delta : -delta;
                d[4] += d[4] >= 0 ? delta : -delta;
                d[8] += d[8] >= 0 ? delta : -delta;
                perturbed++;
                if (!inv3(d, &dinv[(size_t)k * 9], 0.0)) return -4;
            }
        }
        const auto t2 = clk::now();
        // solve, with iterative refinement after perturbed pivots
        std::vector<double> w((size_t)n * 3), r((size_t)n * 3);
        auto apply_inverse = [&](const double *rhs, double *sol) {
            for (int k = 0; k < n; k++)
                for (int e = 0; e < 3; e++) w[3 * k + e] = rhs[3 * an.perm[k] + e];
            for (int i = 0; i < n; i++)
                for (long long p = an.lp[i]; p < an.lp[i + 1]; p++) {
                    const double *l = &lx[(size_t)p * 9];
                    double *t = &w[3 * (size_t)li[p]];
                    const double *s = &w[3 * (size_t)i];
                    for (int e = 0; e < 3; e++) t[e] -= l[3 * e] * s[0] + l[3 * e + 1] * s[1] + l[3 * e + 2] * s[2];
                }
            for (int i = 0; i < n; i++) {
                const double *di = &dinv[(size_t)i * 9];
                double s[3] = {w[3 * i], w[3 * i + 1], w[3 * i + 2]};
                for (int e = 0; e < 3; e++) w[3 * i + e] = di[3 * e] * s[0] + di[3 * e + 1] * s[1] + di[3 * e + 2] * s[2];
            }
            for (int i = n - 1; i >= 0; i--)
                for (long long p = an.lp[i]; p < an.lp[i + 1]; p++) {
                    const double *l = &lx[(size_t)p * 9];
                    const double *s = &w[3 * (size_t)li[p]];
                    double *t = &w[3 * (size_t)i];
                    for (int e = 0; e < 3; e++) t[e] -= l[e] * s[0] + l[3 + e] * s[1] + l[6 + e] * s[2];
                }
            for (int k = 0; k < n; k++)
                for (int e = 0; e < 3; e++) sol[3 * an.perm[k] + e] = w[3 * k + e];
        };
        apply_inverse(b, x);
        for (int step = 0; perturbed && step < 2; step++) {
            for (int i = 0; i < 3 * n; i++) r[i] = -b[i];
            symv(ia, ja, a, n, x, r.data());          // r = A x - b
            std::vector<double> c((size_t)n * 3);
            apply_inverse(r.data(), c.data());
            for (int i = 0; i < 3 * n; i++) x[i] -= c[i];
        }
        if (msglvl > 0) {
            const auto t3 = clk::now();
            auto ms = [](clk::duration d) { return std::chrono::duration<double, std::milli>(d).count(); };
            std::fprintf(stderr,
                         "SolveDouble3 (LDL^T, not PARDISO): %d block rows, %d blocks, L %lld blocks, %d perturbed "
                         "pivots; analysis %.1f ms, factorisation %.1f ms, solve %.1f ms\n",
                         n, ia[n], an.lnz, perturbed, ms(t1 - t0), ms(t2 - t1), ms(t3 - t2));
        }
    } catch (const std::bad_alloc &) {
        g_an = Analysis();
        return -2;
    }
    return 0;
}
