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
//              suits long thin ones (beams); then the elimination tree is postordered;
//   symbolic   elimination tree, column counts, fundamental supernodes merged with their last
//              child where few explicit zeros are added, and each supernode's row structure;
//   numeric    multifrontal block LDL^T: per supernode (in postorder) a dense front holds the
//              matrix's columns and the children's update matrices (extend-add); its first
//              block columns are factorised with 3x3 pivot blocks, NB at a time, and the
//              trailing update matrix F22 - L21 D L21^T is formed by an AVX2/FMA tile kernel
//              (OpenMP across column tiles of large fronts) and kept on a stack for the parent.
//              A pivot block that is singular to working precision gets 1e-8 of the matrix's
//              largest entry added to its diagonal (PARDISO perturbs small pivots by
//              10^-iparm[9] with iparm[9] = 8 in the reference's settings), and the solution then
//              takes up to two steps of iterative refinement;
//   solve      L y = P b, D z = y, L^T w = z, x = P^T w, supernode by supernode.
//
// Analyses are kept for the last 1024 sparsity patterns (the reference re-analyses on every
// call; the result does not depend on this), so calls are reentrant: several threads may
// solve different systems at once.
#include "../../include/icflow_host.h"

#include <immintrin.h>
#include <omp.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <list>
#include <memory>
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

// ---- analysis ---------------------------------------------------------------------------------

struct Analysis {
    int n = -1;
    std::vector<int> ia, ja;       // the pattern analysed
    std::vector<int> perm, iperm;  // new -> old, old -> new (the elimination tree postordered)
    // the permuted matrix's lower part by column: rows i > j of column j, the source block and
    // whether it is transposed; the diagonal's source per column
    std::vector<int> lcp, lri, lsrc;
    std::vector<char> ltr;
    std::vector<int> dsrc;
    // supernodes in postorder: columns [sfirst[s], sfirst[s+1]), rows below them
    // srows[srp[s] .. srp[s+1]) (ascending), children schild[scp[s] .. scp[s+1])
    int ns = 0;
    std::vector<int> sfirst, srows, schild, scp;
    std::vector<long long> srp;
    std::vector<long long> lpos;   // offset of supernode s's dense panel in the factor
    long long lsize = 0;           // doubles of all panels
    long long lblocks = 0;         // off-diagonal 3x3 blocks of L (the panels' block-lower part)
    size_t max_front = 0;          // largest front, doubles
    size_t max_wbuf = 0;           // largest panel, doubles
    size_t stack_peak = 0;         // update-matrix stack, doubles
    double flops = 0;              // of the numeric factorisation
};

// Elimination tree and off-diagonal column counts of the symmetric pattern g under perm.
void etree_counts(const Graph &g, int n, const std::vector<int> &perm, const std::vector<int> &iperm,
                  std::vector<int> &parent, std::vector<int> &cnt)
{
    parent.assign(n, -1);
    cnt.assign(n, 0);
    std::vector<int> flag(n);
    for (int k = 0; k < n; k++) {
        flag[k] = k;
        const int v = perm[k];
        for (int q = g.p[v]; q < g.p[v + 1]; q++) {
            int i = iperm[g.j[q]];
            if (i >= k) continue;
            for (; flag[i] != k; i = parent[i]) {
                if (parent[i] == -1) parent[i] = k;
                cnt[i]++;
                flag[i] = k;
            }
        }
    }
}

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
    std::vector<int> perm0 = factor_blocks(g, n, band) < factor_blocks(g, n, nd.out) ? band : nd.out;
    std::vector<int> iperm0(n);
    for (int k = 0; k < n; k++) iperm0[perm0[k]] = k;
    // postorder the elimination tree (same fill, and every subtree's columns contiguous)
    std::vector<int> parent0, cnt0;
    etree_counts(g, n, perm0, iperm0, parent0, cnt0);
    std::vector<int> post;
    post.reserve(n);
    {
        std::vector<int> head(n, -1), next(n, -1), stk;
        for (int k = n - 1; k >= 0; k--)
            if (parent0[k] >= 0) {
                next[k] = head[parent0[k]];
                head[parent0[k]] = k;
            }
        for (int r = 0; r < n; r++) {
            if (parent0[r] >= 0) continue;
            stk.push_back(r);
            while (!stk.empty()) {
                const int v = stk.back();
                if (head[v] >= 0) {            // descend into the next unvisited child
                    const int c = head[v];
                    head[v] = next[c];
                    stk.push_back(c);
                } else {
                    post.push_back(v);
                    stk.pop_back();
                }
            }
        }
    }
    if ((int)post.size() != n) return -3;
    std::vector<int> ipost(n);
    for (int k = 0; k < n; k++) ipost[post[k]] = k;
    an.perm.resize(n);
    an.iperm.resize(n);
    for (int k = 0; k < n; k++) an.perm[k] = perm0[post[k]];
    for (int k = 0; k < n; k++) an.iperm[an.perm[k]] = k;
    std::vector<int> parent(n), cnt(n), nchild(n, 0);
    for (int k = 0; k < n; k++) {
        const int p0 = parent0[post[k]];
        parent[k] = p0 < 0 ? -1 : ipost[p0];
        cnt[k] = cnt0[post[k]];
        if (parent[k] >= 0) nchild[parent[k]]++;
    }
    // the permuted matrix's lower part by column
    an.lcp.assign(n + 1, 0);
    an.dsrc.assign(n, -1);
    for (int r = 0; r < n; r++)
        for (int q = ia[r]; q < ia[r + 1]; q++) {
            const int pr = an.iperm[r], pc = an.iperm[ja[q]];
            if (pr == pc) an.dsrc[pr] = q;
            else an.lcp[std::min(pr, pc) + 1]++;
        }
    for (int k = 0; k < n; k++) an.lcp[k + 1] += an.lcp[k];
    an.lri.resize(an.lcp[n]);
    an.lsrc.resize(an.lcp[n]);
    an.ltr.resize(an.lcp[n]);
    {
        std::vector<int> at(an.lcp.begin(), an.lcp.end() - 1);
        for (int r = 0; r < n; r++)
            for (int q = ia[r]; q < ia[r + 1]; q++) {
                const int pr = an.iperm[r], pc = an.iperm[ja[q]];
                if (pr == pc) continue;
                const int col = std::min(pr, pc), row = std::max(pr, pc);
                const int a_ = at[col]++;
                an.lri[a_] = row;
                an.lsrc[a_] = q;
                an.ltr[a_] = pr < pc;   // block (r, c) lands above the diagonal: its transpose is below
            }
    }
    for (int k = 0; k < n; k++)
        if (an.dsrc[k] < 0) return -1;   // a diagonal block is missing
    // fundamental supernodes: column j joins j-1's when j is j-1's parent, j-1 its only child,
    // and their structures nest
    std::vector<int> first;
    for (int j = 0; j < n; j++)
        if (!(j > 0 && parent[j - 1] == j && nchild[j] == 1 && cnt[j - 1] == cnt[j] + 1)) first.push_back(j);
    first.push_back(n);
    const int nf = (int)first.size() - 1;
    std::vector<int> col_sn(n);
    for (int s = 0; s < nf; s++)
        for (int j = first[s]; j < first[s + 1]; j++) col_sn[j] = s;
    // relaxed amalgamation: the last child supernode merges into its parent when the merged
    // panel gains few explicit zeros (thresholds in 3x3 block columns: always up to 2, up to 6
    // with < 80% zeros, up to 16 with < 10%, beyond with < 5%)
    struct SN {
        int first, last, nr, orig_parent;
        double zeros;
    };
    std::vector<SN> out;
    out.reserve(nf);
    for (int s = 0; s < nf; s++) {
        const int l = first[s + 1] - 1;
        SN cur{first[s], l, cnt[l], parent[l] < 0 ? -1 : col_sn[parent[l]], 0.0};
        while (!out.empty()) {
            const SN &t = out.back();
            if (t.orig_parent != s || t.last + 1 != cur.first) break;
            const double nct = t.last - t.first + 1, ncs = cur.last - cur.first + 1;
            const double added = nct * (ncs + cur.nr - t.nr);
            const double z = t.zeros + cur.zeros + added;
            const double nc = nct + ncs;
            const double total = nc * (nc + 1) / 2 + nc * cur.nr;
            const double frac = total > 0 ? z / total : 0;
            const bool ok = nc <= 2 || (nc <= 6 && frac < 0.8) || (nc <= 16 && frac < 0.1) || frac < 0.05;
            if (!ok) break;
            cur.first = t.first;
            cur.zeros = z;
            out.pop_back();
        }
        out.push_back(cur);
    }
    an.ns = (int)out.size();
    an.sfirst.resize(an.ns + 1);
    for (int s = 0; s < an.ns; s++) an.sfirst[s] = out[s].first;
    an.sfirst[an.ns] = n;
    for (int s = 0; s < an.ns; s++)
        for (int j = an.sfirst[s]; j < an.sfirst[s + 1]; j++) col_sn[j] = s;
    std::vector<int> sparent(an.ns, -1);
    for (int s = 0; s < an.ns; s++) {
        const int l = an.sfirst[s + 1] - 1;
        sparent[s] = parent[l] < 0 ? -1 : col_sn[parent[l]];
    }
    an.scp.assign(an.ns + 1, 0);
    for (int s = 0; s < an.ns; s++)
        if (sparent[s] >= 0) an.scp[sparent[s] + 1]++;
    for (int s = 0; s < an.ns; s++) an.scp[s + 1] += an.scp[s];
    an.schild.resize(an.scp[an.ns]);
    {
        std::vector<int> at(an.scp.begin(), an.scp.end() - 1);
        for (int s = 0; s < an.ns; s++)
            if (sparent[s] >= 0) an.schild[at[sparent[s]]++] = s;
    }
    // row structures: the matrix's rows below each supernode's columns and its children's rows
    an.srp.assign(an.ns + 1, 0);
    std::vector<int> mark(n, -1), rows;
    for (int s = 0; s < an.ns; s++) {
        const int f = an.sfirst[s], l = an.sfirst[s + 1];
        rows.clear();
        for (int j = f; j < l; j++)
            for (int q = an.lcp[j]; q < an.lcp[j + 1]; q++) {
                const int i = an.lri[q];
                if (i >= l && mark[i] != s) {
                    mark[i] = s;
                    rows.push_back(i);
                }
            }
        for (int q = an.scp[s]; q < an.scp[s + 1]; q++) {
            const int c = an.schild[q];
            for (long long t = an.srp[c]; t < an.srp[c + 1]; t++) {
                const int i = an.srows[t];
                if (i >= l && mark[i] != s) {
                    mark[i] = s;
                    rows.push_back(i);
                }
            }
        }
        std::sort(rows.begin(), rows.end());
        an.srows.insert(an.srows.end(), rows.begin(), rows.end());
        an.srp[s + 1] = (long long)an.srows.size();
    }
    // panel offsets, buffer sizes, the update stack's peak and the flop count
    an.lpos.assign(an.ns + 1, 0);
    size_t sp = 0;
    std::vector<size_t> upd(an.ns, 0);
    for (int s = 0; s < an.ns; s++) {
        const long long nc = an.sfirst[s + 1] - an.sfirst[s], nr = an.srp[s + 1] - an.srp[s];
        const long long M = 3 * (nc + nr), K = 3 * nc;
        an.lpos[s + 1] = an.lpos[s] + M * K;
        an.lblocks += nc * (nc - 1) / 2 + nc * nr;
        an.max_front = std::max(an.max_front, (size_t)(M * M));
        an.max_wbuf = std::max(an.max_wbuf, (size_t)(M * K));
        an.stack_peak = std::max(an.stack_peak, sp + (size_t)(M * M));   // the front above the children's updates
        for (int q = an.scp[s]; q < an.scp[s + 1]; q++) sp -= upd[an.schild[q]];
        upd[s] = (size_t)(9 * nr * nr);
        sp += upd[s];
        an.stack_peak = std::max(an.stack_peak, sp);
        // panel: ~ M K^2 ; trailing update: (M-K)^2 K
        an.flops += (double)M * K * K + (double)(M - K) * (M - K) * K;
    }
    an.lsize = an.lpos[an.ns];
    return 0;
}

bool same_pattern(const Analysis &an, const int *ia, const int *ja, int n)
{
    if (an.n != n || (int)an.ia.size() != n + 1 || an.ia[n] != ia[n]) return false;
    return std::equal(an.ia.begin(), an.ia.end(), ia) && std::equal(an.ja.begin(), an.ja.end(), ja);
}

// Analyses of recent patterns, shared by all threads (each is read-only once built).  Batched
// specimens alternate between their own patterns, so more than one is kept.
struct Cache {
    std::mutex mu;
    std::list<std::pair<uint64_t, std::shared_ptr<const Analysis>>> lru;
    size_t cap = 1024;
};
Cache g_cache;

uint64_t pattern_hash(const int *ia, const int *ja, int n)
{
    uint64_t h = 1469598103934665603ULL ^ (uint64_t)n;
    auto mix = [&](int v) {
        h ^= (uint32_t)v;
        h *= 1099511628211ULL;
    };
    for (int i = 0; i <= n; i++) mix(ia[i]);
    for (int q = 0; q < ia[n]; q++) mix(ja[q]);
    return h;
}

std::shared_ptr<const Analysis> get_analysis(const int *ia, const int *ja, int n, int *err)
{
    const uint64_t h = pattern_hash(ia, ja, n);
    {
        std::lock_guard<std::mutex> lk(g_cache.mu);
        for (auto it = g_cache.lru.begin(); it != g_cache.lru.end(); ++it)
            if (it->first == h && same_pattern(*it->second, ia, ja, n)) {
                g_cache.lru.splice(g_cache.lru.begin(), g_cache.lru, it);
                return g_cache.lru.front().second;
            }
    }
    auto an = std::make_shared<Analysis>();
    if (int r = analyse(*an, ia, ja, n)) {
        *err = r;
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_cache.mu);
    g_cache.lru.emplace_front(h, an);
    while (g_cache.lru.size() > g_cache.cap) g_cache.lru.pop_back();
    return an;
}

// ---- dense kernels ----------------------------------------------------------------------------
// Column-major scalar arrays.  update_lower: C[i, q] -= sum_p A[i, p] * B[q, p] for the
// columns q in [q0, q1) and the rows i from the first row of q's 3x3 block row to rows - 1
// (the block-lower part), p in [p0, p1).  B is packed 4 columns at a time; tiles of 12 (or 8)
// rows x 4 columns read A in place.  The tiles also write some entries of blocks above the
// diagonal, which nothing reads.

std::atomic<int> g_threads{0};   // threads for the large updates (0: OpenMP's default)

inline void pack_b(const double *B, int ldb, int q, int nq, int p0, int p1, double *bp)
{
    for (int p = p0; p < p1; p++)
        for (int jj = 0; jj < 4; jj++) bp[4 * (p - p0) + jj] = jj < nq ? B[q + jj + (size_t)ldb * p] : 0.0;
}

__attribute__((target("avx2,fma"))) void tile8x4_avx2(double *C, int ldc, const double *A, int lda, const double *bp,
                                                      int K)
{
    __m256d c00 = _mm256_setzero_pd(), c01 = c00, c02 = c00, c03 = c00, c10 = c00, c11 = c00, c12 = c00, c13 = c00;
    for (int p = 0; p < K; p++) {
        const double *a = A + (size_t)lda * p;
        const __m256d a0 = _mm256_loadu_pd(a), a1 = _mm256_loadu_pd(a + 4);
        const double *b = bp + 4 * p;
        __m256d b0 = _mm256_broadcast_sd(b), b1 = _mm256_broadcast_sd(b + 1);
        c00 = _mm256_fmadd_pd(a0, b0, c00);
        c10 = _mm256_fmadd_pd(a1, b0, c10);
        c01 = _mm256_fmadd_pd(a0, b1, c01);
        c11 = _mm256_fmadd_pd(a1, b1, c11);
        b0 = _mm256_broadcast_sd(b + 2);
        b1 = _mm256_broadcast_sd(b + 3);
        c02 = _mm256_fmadd_pd(a0, b0, c02);
        c12 = _mm256_fmadd_pd(a1, b0, c12);
        c03 = _mm256_fmadd_pd(a0, b1, c03);
        c13 = _mm256_fmadd_pd(a1, b1, c13);
    }
    double *c = C;
    _mm256_storeu_pd(c, _mm256_sub_pd(_mm256_loadu_pd(c), c00));
    _mm256_storeu_pd(c + 4, _mm256_sub_pd(_mm256_loadu_pd(c + 4), c10));
    c += ldc;
    _mm256_storeu_pd(c, _mm256_sub_pd(_mm256_loadu_pd(c), c01));
    _mm256_storeu_pd(c + 4, _mm256_sub_pd(_mm256_loadu_pd(c + 4), c11));
    c += ldc;
    _mm256_storeu_pd(c, _mm256_sub_pd(_mm256_loadu_pd(c), c02));
    _mm256_storeu_pd(c + 4, _mm256_sub_pd(_mm256_loadu_pd(c + 4), c12));
    c += ldc;
    _mm256_storeu_pd(c, _mm256_sub_pd(_mm256_loadu_pd(c), c03));
    _mm256_storeu_pd(c + 4, _mm256_sub_pd(_mm256_loadu_pd(c + 4), c13));
}

__attribute__((target("avx2,fma"))) void tile12x4_avx2(double *C, int ldc, const double *A, int lda, const double *bp,
                                                       int K)
{
    __m256d c00 = _mm256_setzero_pd(), c01 = c00, c02 = c00, c03 = c00, c10 = c00, c11 = c00, c12 = c00, c13 = c00;
    __m256d c20 = c00, c21 = c00, c22 = c00, c23 = c00;
    for (int p = 0; p < K; p++) {
        const double *a = A + (size_t)lda * p;
        const __m256d a0 = _mm256_loadu_pd(a), a1 = _mm256_loadu_pd(a + 4), a2 = _mm256_loadu_pd(a + 8);
        const double *b = bp + 4 * p;
        __m256d bb = _mm256_broadcast_sd(b);
        c00 = _mm256_fmadd_pd(a0, bb, c00);
        c10 = _mm256_fmadd_pd(a1, bb, c10);
        c20 = _mm256_fmadd_pd(a2, bb, c20);
        bb = _mm256_broadcast_sd(b + 1);
        c01 = _mm256_fmadd_pd(a0, bb, c01);
        c11 = _mm256_fmadd_pd(a1, bb, c11);
        c21 = _mm256_fmadd_pd(a2, bb, c21);
        bb = _mm256_broadcast_sd(b + 2);
        c02 = _mm256_fmadd_pd(a0, bb, c02);
        c12 = _mm256_fmadd_pd(a1, bb, c12);
        c22 = _mm256_fmadd_pd(a2, bb, c22);
        bb = _mm256_broadcast_sd(b + 3);
        c03 = _mm256_fmadd_pd(a0, bb, c03);
        c13 = _mm256_fmadd_pd(a1, bb, c13);
        c23 = _mm256_fmadd_pd(a2, bb, c23);
    }
    const __m256d acc[4][3] = {{c00, c10, c20}, {c01, c11, c21}, {c02, c12, c22}, {c03, c13, c23}};
    for (int jj = 0; jj < 4; jj++) {
        double *c = C + (size_t)ldc * jj;
        for (int v = 0; v < 3; v++) _mm256_storeu_pd(c + 4 * v, _mm256_sub_pd(_mm256_loadu_pd(c + 4 * v), acc[jj][v]));
    }
}

// any tile: ni rows x nq columns
void tile_any(double *C, int ldc, const double *A, int lda, const double *bp, int K, int ni, int nq)
{
    double acc[8][4] = {};
    for (int p = 0; p < K; p++) {
        const double *a = A + (size_t)lda * p;
        const double *b = bp + 4 * p;
        for (int ii = 0; ii < ni; ii++)
            for (int jj = 0; jj < 4; jj++) acc[ii][jj] += a[ii] * b[jj];
    }
    for (int jj = 0; jj < nq; jj++)
        for (int ii = 0; ii < ni; ii++) C[ii + (size_t)ldc * jj] -= acc[ii][jj];
}

const bool g_avx2 = __builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma");

void update_lower(double *C, int ldc, int rows, int q0, int q1, const double *A, int lda, const double *B, int ldb,
                  int p0, int p1)
{
    const int K = p1 - p0;
    if (K <= 0 || q1 <= q0) return;
    const int ntiles = (q1 - q0 + 3) / 4;
    const double work = (double)(rows - q0) * (q1 - q0) * K;
    const int thr = g_threads.load();
    auto body = [&](int t, double *bp) {
        const int q = q0 + 4 * t, nq = std::min(4, q1 - q);
        pack_b(B, ldb, q, nq, p0, p1, bp);
        int i = 3 * (q / 3);
        const double *Ap = A + (size_t)lda * p0;
        double *Cq = C + (size_t)ldc * q;
        if (g_avx2 && nq == 4) {
            for (; i + 12 <= rows; i += 12) tile12x4_avx2(Cq + i, ldc, Ap + i, lda, bp, K);
            for (; i + 8 <= rows; i += 8) tile8x4_avx2(Cq + i, ldc, Ap + i, lda, bp, K);
        }
        for (; i < rows; i += 8) tile_any(Cq + i, ldc, Ap + i, lda, bp, K, std::min(8, rows - i), nq);
    };
    if (work > 2e7 && thr != 1 && omp_get_max_threads() > 1 && !omp_in_parallel()) {
#pragma omp parallel num_threads(thr > 0 ? thr : omp_get_max_threads())
        {
            std::vector<double> bp((size_t)4 * K);
#pragma omp for schedule(dynamic, 1)
            for (int t = 0; t < ntiles; t++) body(t, bp.data());
        }
    } else {
        std::vector<double> bp((size_t)4 * K);
        for (int t = 0; t < ntiles; t++) body(t, bp.data());
    }
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

// ---- numeric factorisation (multifrontal) and solve --------------------------------------------

struct Factor {
    std::vector<double> L;      // the supernodes' dense panels (M x 3nc, column-major)
    std::vector<double> dinv;   // inverse pivot blocks, 3x3 row-major per block column
    int perturbed = 0;
};

constexpr int NB = 32;   // block columns factorised together before the rest of a panel is updated

int factorise(const Analysis &an, const double *a, Factor &fx, int msglvl = 0)
{
    const int n = an.n;
    // msglvl > 1: a progress line at every tenth of the factorisation's flops
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    double done = 0;
    int tenth = 1;
    double amax = 0;
    for (long long i = 0; i < (long long)an.ia[n] * 9; i++) amax = std::max(amax, std::fabs(a[i]));
    const double delta = 1e-8 * (amax > 0 ? amax : 1.0);
    fx.L.assign((size_t)an.lsize, 0.0);
    fx.dinv.assign((size_t)n * 9, 0.0);
    fx.perturbed = 0;
    std::vector<double> stack(an.stack_peak), wbuf(an.max_wbuf);
    std::vector<size_t> upd(an.ns, 0);
    std::vector<int> pos(n, -1);
    size_t sp = 0;
    for (int s = 0; s < an.ns; s++) {
        const int f = an.sfirst[s], nc = an.sfirst[s + 1] - f;
        const int nr = (int)(an.srp[s + 1] - an.srp[s]);
        const int *rows = an.srows.data() + an.srp[s];
        const int m = nc + nr, M = 3 * m, K = 3 * nc;
        // children's updates are on top of the stack; the front goes above them
        size_t base = sp;
        for (int q = an.scp[s]; q < an.scp[s + 1]; q++) base -= upd[an.schild[q]];
        double *F = stack.data() + sp;
        for (int q = 0; q < M; q++) {
            const int r0 = 3 * (q / 3);
            std::memset(F + (size_t)M * q + r0, 0, sizeof(double) * (M - r0));
        }
        for (int t = 0; t < nc; t++) pos[f + t] = t;
        for (int t = 0; t < nr; t++) pos[rows[t]] = nc + t;
        // the matrix's columns
        for (int t = 0; t < nc; t++) {
            const int j = f + t;
            const double *d = a + (size_t)an.dsrc[j] * 9;
            for (int x = 0; x < 3; x++)
                for (int y = 0; y < 3; y++) F[(3 * t + x) + (size_t)M * (3 * t + y)] += d[3 * x + y];
            for (int q = an.lcp[j]; q < an.lcp[j + 1]; q++) {
                const int ti = pos[an.lri[q]];
                const double *b = a + (size_t)an.lsrc[q] * 9;
                const bool tr = an.ltr[q];
                for (int x = 0; x < 3; x++)
                    for (int y = 0; y < 3; y++)
                        F[(3 * ti + x) + (size_t)M * (3 * t + y)] += tr ? b[3 * y + x] : b[3 * x + y];
            }
        }
        // extend-add of the children's update matrices (block-lower part)
        {
            size_t off = base;
            for (int q = an.scp[s]; q < an.scp[s + 1]; q++) {
                const int c = an.schild[q];
                const int ncr = (int)(an.srp[c + 1] - an.srp[c]);
                const int *crow = an.srows.data() + an.srp[c];
                const int Mc = 3 * ncr;
                const double *U = stack.data() + off;
                for (int bj = 0; bj < ncr; bj++) {
                    const int tj = pos[crow[bj]];
                    for (int y = 0; y < 3; y++) {
                        const double *u = U + (size_t)Mc * (3 * bj + y);
                        double *fc = F + (size_t)M * (3 * tj + y);
                        for (int bi = bj; bi < ncr; bi++) {
                            const int ti = pos[crow[bi]];
                            fc[3 * ti] += u[3 * bi];
                            fc[3 * ti + 1] += u[3 * bi + 1];
                            fc[3 * ti + 2] += u[3 * bi + 2];
                        }
                    }
                }
                off += upd[c];
            }
        }
        // partial factorisation of the first nc block columns, NB at a time: W = L D is kept
        // beside L for the updates
        double *W = wbuf.data();
        for (int c0 = 0; c0 < nc; c0 += NB) {
            const int c1 = std::min(nc, c0 + NB);
            for (int c = c0; c < c1; c++) {
                double d[9], di[9];
                for (int x = 0; x < 3; x++)
                    for (int y = 0; y < 3; y++) d[3 * x + y] = F[(3 * c + x) + (size_t)M * (3 * c + y)];
                if (!inv3(d, di, 1e-13)) {
                    d[0] += d[0] >= 0 ? delta : -delta;
                    d[4] += d[4] >= 0 ? delta : -delta;
                    d[8] += d[8] >= 0 ? delta : -delta;
                    fx.perturbed++;
                    if (!inv3(d, di, 0.0)) return -4;
                }
                std::memcpy(&fx.dinv[(size_t)(f + c) * 9], di, sizeof di);
                // L(r, c) = F(r, c) D^-1 for the rows below the pivot block; W keeps F(r, c)
                double *f0 = F + (size_t)M * (3 * c), *f1 = f0 + M, *f2 = f1 + M;
                double *w0 = W + (size_t)M * (3 * c), *w1 = w0 + M, *w2 = w1 + M;
                for (int i = 3 * (c + 1); i < M; i++) {
                    const double x0 = f0[i], x1 = f1[i], x2 = f2[i];
                    w0[i] = x0;
                    w1[i] = x1;
                    w2[i] = x2;
                    f0[i] = x0 * di[0] + x1 * di[3] + x2 * di[6];
                    f1[i] = x0 * di[1] + x1 * di[4] + x2 * di[7];
                    f2[i] = x0 * di[2] + x1 * di[5] + x2 * di[8];
                }
                // rank-3 update of the chunk's remaining columns
                for (int q = 3 * (c + 1); q < 3 * c1; q++) {
                    const double y0 = w0[q], y1 = w1[q], y2 = w2[q];
                    double *fq = F + (size_t)M * q;
                    for (int i = 3 * (q / 3); i < M; i++) fq[i] -= f0[i] * y0 + f1[i] * y1 + f2[i] * y2;
                }
            }
            // the rest of the panel from this chunk
            update_lower(F, M, M, 3 * c1, K, F, M, W, M, 3 * c0, 3 * c1);
        }
        // the trailing update matrix: F22 -= L21 W21^T
        update_lower(F, M, M, K, M, F, M, W, M, 0, K);
        // keep the panel; move the update matrix down over the children's
        std::memcpy(fx.L.data() + an.lpos[s], F, sizeof(double) * (size_t)M * K);
        const int Mu = 3 * nr;
        double *U = stack.data() + base;
        for (int y = 0; y < Mu; y++)
            std::memmove(U + (size_t)Mu * y, F + (size_t)M * (K + y) + K, sizeof(double) * Mu);
        upd[s] = (size_t)Mu * Mu;
        sp = base + upd[s];
        if (msglvl > 1) {
            done += (double)M * K * K + (double)(M - K) * (M - K) * K;
            for (; tenth <= 10 && done >= an.flops * tenth / 10; tenth++)
                std::fprintf(stderr, "SolveDouble3: factorisation %d%% (%.1f s)\n", 10 * tenth,
                             std::chrono::duration<double>(clk::now() - t0).count());
        }
    }
    return 0;
}

// x = A^-1 b through the factor (b and x in the original numbering, 3n values)
void apply_inverse(const Analysis &an, const Factor &fx, const double *b, double *x, std::vector<double> &w)
{
    const int n = an.n;
    w.resize((size_t)3 * n);
    for (int k = 0; k < n; k++)
        for (int e = 0; e < 3; e++) w[3 * k + e] = b[3 * an.perm[k] + e];
    for (int s = 0; s < an.ns; s++) {
        const int f = an.sfirst[s], nc = an.sfirst[s + 1] - f;
        const int nr = (int)(an.srp[s + 1] - an.srp[s]);
        const int *rows = an.srows.data() + an.srp[s];
        const int M = 3 * (nc + nr), K = 3 * nc;
        const double *Lp = fx.L.data() + an.lpos[s];
        double *ws = w.data() + 3 * f;
        for (int p = 0; p < K; p++) {
            const double xp = ws[p];
            const double *l = Lp + (size_t)M * p;
            for (int i = 3 * (p / 3 + 1); i < K; i++) ws[i] -= l[i] * xp;
            for (int t = 0; t < nr; t++) {
                double *wr = w.data() + 3 * rows[t];
                wr[0] -= l[K + 3 * t] * xp;
                wr[1] -= l[K + 3 * t + 1] * xp;
                wr[2] -= l[K + 3 * t + 2] * xp;
            }
        }
    }
    for (int k = 0; k < n; k++) {
        const double *di = &fx.dinv[(size_t)k * 9];
        const double s0 = w[3 * k], s1 = w[3 * k + 1], s2 = w[3 * k + 2];
        for (int e = 0; e < 3; e++) w[3 * k + e] = di[3 * e] * s0 + di[3 * e + 1] * s1 + di[3 * e + 2] * s2;
    }
    for (int s = an.ns - 1; s >= 0; s--) {
        const int f = an.sfirst[s], nc = an.sfirst[s + 1] - f;
        const int nr = (int)(an.srp[s + 1] - an.srp[s]);
        const int *rows = an.srows.data() + an.srp[s];
        const int M = 3 * (nc + nr), K = 3 * nc;
        const double *Lp = fx.L.data() + an.lpos[s];
        double *ws = w.data() + 3 * f;
        for (int p = 0; p < K; p++) {
            const double *l = Lp + (size_t)M * p + K;
            double acc = 0;
            for (int t = 0; t < nr; t++) {
                const double *wr = w.data() + 3 * rows[t];
                acc += l[3 * t] * wr[0] + l[3 * t + 1] * wr[1] + l[3 * t + 2] * wr[2];
            }
            ws[p] -= acc;
        }
        for (int p = K - 1; p >= 0; p--) {
            const double *l = Lp + (size_t)M * p;
            double acc = 0;
            for (int i = 3 * (p / 3 + 1); i < K; i++) acc += l[i] * ws[i];
            ws[p] -= acc;
        }
    }
    for (int k = 0; k < n; k++)
        for (int e = 0; e < 3; e++) x[3 * an.perm[k] + e] = w[3 * k + e];
}

}  // namespace

extern "C" int icfh_solve_threads(int threads)
{
    if (threads < 0) return -1;
    g_threads.store(threads);
    return 0;
}

extern "C" int icfh_solve_analyse(const int *ja, const int *ia, int n, long long *factor_doubles,
                                  long long *peak_doubles, double *flops)
{
    if (n <= 0 || !ia || !ja || !factor_doubles || !peak_doubles || !flops) return -1;
    try {
        int err = 0;
        std::shared_ptr<const Analysis> an = get_analysis(ia, ja, n, &err);
        if (!an) return err;
        *factor_doubles = an->lsize;
        *peak_doubles = an->lsize + (long long)an->stack_peak + (long long)an->max_front + (long long)an->max_wbuf;
        *flops = an->flops;
    } catch (const std::bad_alloc &) {
        return -2;
    }
    return 0;
}

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
    try {
        int err = 0;
        std::shared_ptr<const Analysis> anp = get_analysis(ia, ja, n, &err);
        if (!anp) return err;
        const Analysis &an = *anp;
        const auto t1 = clk::now();
        Factor fx;
        if (int r = factorise(an, a, fx, msglvl)) return r;
        const auto t2 = clk::now();
        // solve, with iterative refinement after perturbed pivots
        std::vector<double> w, r((size_t)n * 3), c((size_t)n * 3);
        apply_inverse(an, fx, b, x, w);
        for (int step = 0; fx.perturbed && step < 2; step++) {
            for (int i = 0; i < 3 * n; i++) r[i] = -b[i];
            symv(ia, ja, a, n, x, r.data());          // r = A x - b
            apply_inverse(an, fx, r.data(), c.data(), w);
            for (int i = 0; i < 3 * n; i++) x[i] -= c[i];
        }
        if (msglvl > 0) {
            const auto t3 = clk::now();
            auto ms = [](clk::duration d) { return std::chrono::duration<double, std::milli>(d).count(); };
            std::fprintf(stderr,
                         "SolveDouble3 (LDL^T, not PARDISO): %d block rows, %d blocks, L %lld blocks in %d supernodes, "
                         "%.3g GFLOP, %d perturbed pivots; analysis %.1f ms, factorisation %.1f ms (%.2f GFLOP/s), "
                         "solve %.1f ms\n",
                         n, ia[n], an.lblocks, an.ns, an.flops * 1e-9, fx.perturbed, ms(t1 - t0), ms(t2 - t1),
                         an.flops * 1e-9 / std::max(1e-9, ms(t2 - t1) * 1e-3), ms(t3 - t2));
        }
    } catch (const std::bad_alloc &) {
        return -2;
    }
    return 0;
}
