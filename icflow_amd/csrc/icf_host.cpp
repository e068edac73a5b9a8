// icf_host.cpp -- host-side helpers of libicflow_hip.so (include/icflow_host.h): model
// constants, BCSR pattern + slot lookup, pcsr packing and the narrow-phase tuple
// de-duplication, restated from icFlow's C# host.
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <algorithm>
#include <unordered_set>
#include <vector>

#include "../../include/icflow_host.h"

struct icfh_csr {
    int N = 0, nnz = 0;
    std::vector<int> rows, cols;
};

namespace {

double macaulay(double a, double b) { return a > b ? a - b : 0; }

// Slot of (row, col) by binary search in the row's sorted columns.
inline int lookup(const icfh_csr *c, int row, int col)
{
    if (row < 0 || row >= c->N) return -1;
    const int *b = c->cols.data() + c->rows[row], *e = c->cols.data() + c->rows[row + 1];
    const int *p = std::lower_bound(b, e, col);
    return (p != e && *p == col) ? (int)(p - c->cols.data()) : -1;
}

// The reference looks pairs up in a Dictionary that throws on a missing key; a missing pair
// here means the pattern was built from different units, reported as ICF_E_INDEX.
template <int K>
int pack_unit(const icfh_csr *c, const int *alt_of_unit_nodes, int *slots /* K*K */, int *rows /* K */)
{
    for (int i = 0; i < K; i++) {
        const int ai = alt_of_unit_nodes[i];
        for (int j = 0; j < K; j++) {
            const int aj = alt_of_unit_nodes[j];
            int s = -1;
            if (ai >= 0 && aj >= 0 && aj >= ai) {
                s = lookup(c, ai, aj);
                if (s < 0) return ICF_E_INDEX;
            }
            slots[i * K + j] = s;
        }
        rows[i] = ai >= 0 ? ai : -1;
    }
    return 0;
}

}  // namespace

extern "C" {

int icfh_compute_constants(const icf_model_prms *p, icf_consts *o)
{
    if (!p || !o) return ICF_E_INVALID;
    memset(o, 0, sizeof(*o));
    // consistent mass per unit volume, ModelPrms.cs:239-248
    const double cm = p->rho / 20.0;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            for (int m = 0; m < 3; m++) {
                int col = i * 3 + m, row = j * 3 + m;
                o->M[col + 12 * row] = (col == row) ? 2 * cm : cm;
            }
    // isotropic elasticity, ModelPrms.cs:250-257
    const double c1 = p->Y / ((1.0 + p->nu) * (1.0 - 2.0 * p->nu));
    double E3[6][6] = {{0}};
    E3[0][0] = E3[1][1] = E3[2][2] = (1.0 - p->nu) * c1;
    E3[0][1] = E3[0][2] = E3[1][2] = E3[1][0] = E3[2][0] = E3[2][1] = p->nu * c1;
    E3[3][3] = E3[4][4] = E3[5][5] = (0.5 - p->nu) * c1;
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) o->E[i + 6 * j] = E3[i][j];
    o->NewmarkBeta = p->NewmarkBeta;
    o->NewmarkGamma = p->NewmarkGamma;
    o->dampingMass = p->dampingMass;
    o->dampingStiffness = p->dampingStiffness;
    o->rho = p->rho;
    o->gravity = p->gravity;
    o->YoungsModulus = p->Y;
    // PPR derived constants, ModelPrms.cs:259-285
    o->G_fn = p->phi_n;
    o->G_ft = p->phi_t;
    o->f_tn = p->sigma_max;
    o->f_tt = p->tau_max;
    o->alpha = p->alpha;
    o->beta = p->beta;
    o->rn = p->lambda_n;
    o->rt = p->lambda_t;
    o->pMnt = macaulay(o->G_fn, o->G_ft);
    o->pMtn = macaulay(o->G_ft, o->G_fn);
    const double rn2 = o->rn * o->rn, rt2 = o->rt * o->rt;
    const double a = o->alpha, b = o->beta;
    o->p_m = (a * (a - 1.0) * rn2) / (1.0 - a * rn2);
    o->p_n = (b * (b - 1.0) * rt2) / (1.0 - b * rt2);
    if (o->G_fn < o->G_ft) {
        o->gam_n = pow(a / o->p_m, o->p_m);
        o->gam_t = -o->G_ft * pow(b / o->p_n, o->p_n);
    } else {
        o->gam_n = -o->G_fn * pow(a / o->p_m, o->p_m);
        o->gam_t = pow(b / o->p_n, o->p_n);
    }
    o->deln = (o->G_fn / o->f_tn) * a * o->rn * pow((1.0 - o->rn), (a - 1.0)) * ((a / o->p_m) + 1.0) *
              pow(((a / o->p_m) * o->rn + 1.0), (o->p_m - 1.0));
    o->delt = (o->G_ft / o->f_tt) * b * o->rt * pow((1.0 - o->rt), (b - 1.0)) * ((b / o->p_n) + 1.0) *
              pow(((b / o->p_n) * o->rt + 1.0), (o->p_n - 1.0));
    // Gauss-point shape functions, ModelPrms.cs:288-307: (1/6,2/3), (2/3,1/6), (1/6,1/6)
    const double gp[3][2] = {{1.0 / 6.0, 2.0 / 3.0}, {2.0 / 3.0, 1.0 / 6.0}, {1.0 / 6.0, 1.0 / 6.0}};
    double sf[3][3];
    for (int g = 0; g < 3; g++) {
        sf[0][g] = 1.0 - gp[g][0] - gp[g][1];
        sf[1][g] = gp[g][0];
        sf[2][g] = gp[g][1];
    }
    for (int i = 0; i < 3; i++)
        for (int g = 0; g < 3; g++) o->sf[i * 3 + g] = sf[i][g];
    // B[gp][3][18] = [s (x) I3, -s (x) I3], ModelPrms.cs:324-349
    for (int g = 0; g < 3; g++)
        for (int n = 0; n < 3; n++)
            for (int k = 0; k < 3; k++) {
                o->B[g * 54 + k * 18 + 3 * n + k] = sf[n][g];
                o->B[g * 54 + k * 18 + 9 + 3 * n + k] = -sf[n][g];
            }
    return 0;
}

int icfh_csr_create(icfh_csr **out)
{
    if (!out) return ICF_E_INVALID;
    *out = new icfh_csr();
    return 0;
}

int icfh_csr_destroy(icfh_csr *c)
{
    delete c;
    return 0;
}

int icfh_csr_build(icfh_csr *c, const int *alt_elem, int nElems, const int *alt_cz, int nCZs, const int *alt_imp,
                   int nImpacts)
{
    if (!c || nElems < 0 || nCZs < 0 || nImpacts < 0) return ICF_E_INVALID;
    if ((nElems && !alt_elem) || (nCZs && !alt_cz) || (nImpacts && !alt_imp)) return ICF_E_INVALID;
    // N = largest static row + 1 (CSRDictionary.updateMaxRowIndex via AddStatic)
    int maxrow = -1;
    for (long long i = 0; i < 4LL * nElems; i++) maxrow = std::max(maxrow, alt_elem[i]);
    for (long long i = 0; i < 6LL * nCZs; i++) maxrow = std::max(maxrow, alt_cz[i]);
    const int N = maxrow + 1;
    // bucket upper pairs by row (counting sort), then sort + unique per row
    std::vector<int> cnt((size_t)N + 1, 0);
    auto visit = [&](auto &&emit) {
        for (int u = 0; u < nCZs; u++) {          // static: CZs then elements (MeshCollection.cs:99-108)
            const int *a = alt_cz + 6LL * u;
            for (int i = 0; i < 6; i++)
                for (int j = 0; j < 6; j++)
                    if (a[i] >= 0 && a[j] >= 0 && a[j] >= a[i]) emit(a[i], a[j]);
        }
        for (int u = 0; u < nElems; u++) {
            const int *a = alt_elem + 4LL * u;
            for (int i = 0; i < 4; i++)
                for (int j = 0; j < 4; j++)
                    if (a[i] >= 0 && a[j] >= 0 && a[j] >= a[i]) emit(a[i], a[j]);
        }
        for (int u = 0; u < nImpacts; u++) {      // dynamic: strictly upper (ImplicitModel3.cs:190)
            const int *a = alt_imp + 4LL * u;
            for (int i = 0; i < 4; i++)
                for (int j = 0; j < 4; j++)
                    if (a[i] >= 0 && a[j] >= 0 && a[j] > a[i] && a[i] < N) emit(a[i], a[j]);
        }
    };
    int bad = 0;
    visit([&](int r, int col) {
        if (r < N) cnt[r + 1]++;
        else bad = 1;
        (void)col;
    });
    if (bad) return ICF_E_INDEX;
    for (int i = 0; i < N; i++) cnt[i + 1] += cnt[i];
    std::vector<int> tmp(cnt[N]);
    std::vector<int> cur(cnt.begin(), cnt.end() - 1);
    visit([&](int r, int col) { tmp[cur[r]++] = col; });
    c->N = N;
    c->rows.assign((size_t)N + 1, 0);
    c->cols.clear();
    c->cols.reserve(tmp.size() / 3 + 16);
    for (int r = 0; r < N; r++) {
        int *b = tmp.data() + cnt[r], *e = tmp.data() + cnt[r + 1];
        std::sort(b, e);
        e = std::unique(b, e);
        c->rows[r] = (int)c->cols.size();
        c->cols.insert(c->cols.end(), b, e);
    }
    c->rows[N] = (int)c->cols.size();
    c->nnz = (int)c->cols.size();
    return 0;
}

int icfh_csr_dims(const icfh_csr *c, int *N, int *nnz)
{
    if (!c) return ICF_E_INVALID;
    if (N) *N = c->N;
    if (nnz) *nnz = c->nnz;
    return 0;
}

int icfh_csr_copy(const icfh_csr *c, int *rows, int *cols)
{
    if (!c) return ICF_E_INVALID;
    if (rows) memcpy(rows, c->rows.data(), sizeof(int) * ((size_t)c->N + 1));
    if (cols && c->nnz) memcpy(cols, c->cols.data(), sizeof(int) * (size_t)c->nnz);
    return 0;
}

int icfh_csr_lookup(const icfh_csr *c, int row, int col) { return c ? lookup(c, row, col) : -1; }

int icfh_pack_elems(const icfh_csr *c, const int *nodes, const int *alt, int nElems, int es, int *ie)
{
    if (!c || nElems < 0 || es < nElems || (nElems && (!nodes || !alt || !ie))) return ICF_E_INVALID;
    for (int u = 0; u < nElems; u++) {
        int a[4], slots[16], rows[4];
        for (int i = 0; i < 4; i++) a[i] = alt[nodes[4LL * u + i]];
        if (int r = pack_unit<4>(c, a, slots, rows)) return r;
        for (int i = 0; i < 4; i++) {
            ie[u + (size_t)es * i] = nodes[4LL * u + i];
            ie[u + (size_t)es * (20 + i)] = rows[i];
        }
        for (int q = 0; q < 16; q++) ie[u + (size_t)es * (4 + q)] = slots[q];
    }
    return 0;
}

int icfh_pack_czs(const icfh_csr *c, const int *nodes, const int *alt, int nCZs, int cs, int *icz)
{
    if (!c || nCZs < 0 || cs < nCZs || (nCZs && (!nodes || !alt || !icz))) return ICF_E_INVALID;
    for (int u = 0; u < nCZs; u++) {
        int a[6], slots[36], rows[6];
        for (int i = 0; i < 6; i++) a[i] = alt[nodes[6LL * u + i]];
        if (int r = pack_unit<6>(c, a, slots, rows)) return r;
        icz[u + (size_t)cs * 0] = 0;   // current failed (GPU_Functionality.cs:384-386)
        icz[u + (size_t)cs * 2] = 0;
        icz[u + (size_t)cs * 3] = 0;
        for (int i = 0; i < 6; i++) {
            icz[u + (size_t)cs * (4 + i)] = nodes[6LL * u + i];
            icz[u + (size_t)cs * (46 + i)] = rows[i];
        }
        for (int q = 0; q < 36; q++) icz[u + (size_t)cs * (10 + q)] = slots[q];
    }
    return 0;
}

int icfh_pack_impacts(const icfh_csr *c, const int *nl2, int ts, const int *alt, int nImpacts, int cr, int *icr)
{
    if (!c || nImpacts < 0 || ts < nImpacts || cr < nImpacts || (nImpacts && (!nl2 || !alt || !icr)))
        return ICF_E_INVALID;
    for (int u = 0; u < nImpacts; u++) {
        int nd[4], a[4], slots[16], rows[4];
        for (int i = 0; i < 4; i++) {
            nd[i] = nl2[u + (size_t)ts * i];
            a[i] = alt[nd[i]];
        }
        if (int r = pack_unit<4>(c, a, slots, rows)) return r;
        for (int i = 0; i < 4; i++) {
            icr[u + (size_t)cr * i] = nd[i];
            icr[u + (size_t)cr * (20 + i)] = rows[i];
        }
        for (int q = 0; q < 16; q++) icr[u + (size_t)cr * (4 + q)] = slots[q];
    }
    return 0;
}

int icfh_nl2set(const int *masks, const int *pairs, int nPairs, const int *tet_nodes, int es, int *out, int os,
                int max_out, int *nOut)
{
    if (nPairs < 0 || !nOut || (nPairs && (!masks || !pairs || !tet_nodes)) || max_out < 0 || os < max_out ||
        (max_out && !out))
        return ICF_E_INVALID;
    std::unordered_set<uint64_t> seen;
    seen.reserve((size_t)nPairs + 16);
    int n = 0;
    auto add = [&](int node, int tet) -> int {
        uint64_t key = ((uint64_t)(uint32_t)node << 32) | (uint32_t)tet;
        if (!seen.insert(key).second) return 0;
        if (n >= max_out) return ICF_E_INDEX;
        out[n] = node;
        out[n + (size_t)os] = tet;
        n++;
        return 0;
    };
    for (int i = 0; i < nPairs; i++) {
        const int bits = masks[2LL * i];
        if (!bits) continue;
        const int tA = pairs[2LL * i], tB = pairs[2LL * i + 1];
        for (int k = 0; k < 4; k++)   // vertices of B inside A
            if (bits & (1 << k))
                if (int r = add(tet_nodes[tB + (size_t)es * k], tA)) return r;
        for (int k = 0; k < 4; k++)   // vertices of A inside B
            if (bits & (16 << k))
                if (int r = add(tet_nodes[tA + (size_t)es * k], tB)) return r;
    }
    *nOut = n;
    return 0;
}

int icfh_kdop_pairs(const double *dn, int ns, const int *ie, int es, int nT, int exclude_shared, int *pairs,
                    long long max_pairs, long long *nOut)
{
    if (!dn || !ie || !nOut || nT < 0 || es < nT || max_pairs < 0 || (max_pairs && !pairs)) return ICF_E_INVALID;
    *nOut = 0;
    if (nT == 0) return 0;
    // k-DOP bounds (kDOP24.Expand): x, y, z, x+y, x+z, y+z, x-y, x-z, y-z, x+y-z, x+z-y, y+z-x
    std::vector<double> lo((size_t)nT * 12), hi((size_t)nT * 12);
    double ext = 0, gmin[3] = {1e300, 1e300, 1e300};
    for (int t = 0; t < nT; t++) {
        double *l = &lo[(size_t)t * 12], *h = &hi[(size_t)t * 12];
        for (int d = 0; d < 12; d++) { l[d] = 1e300; h[d] = -1e300; }
        for (int v = 0; v < 4; v++) {
            int n = ie[t + (size_t)es * v];
            double x = dn[n + (size_t)ns * 15], y = dn[n + (size_t)ns * 16], z = dn[n + (size_t)ns * 17];
            double p[12] = {x, y, z, x + y, x + z, y + z, x - y, x - z, y - z, x + y - z, x + z - y, y + z - x};
            for (int d = 0; d < 12; d++) { l[d] = std::min(l[d], p[d]); h[d] = std::max(h[d], p[d]); }
        }
        for (int k = 0; k < 3; k++) { ext = std::max(ext, h[k] - l[k]); gmin[k] = std::min(gmin[k], l[k]); }
    }
    const double cs = ext > 0 ? ext * 1.000001 : 1.0;
    // bin by the cell of the AABB's min corner; overlapping boxes differ by <= 1 cell per axis
    std::vector<long long> cell((size_t)nT);
    int dim[3] = {1, 1, 1};
    std::vector<int> ci((size_t)nT * 3);
    for (int t = 0; t < nT; t++)
        for (int k = 0; k < 3; k++) {
            int c = (int)floor((lo[(size_t)t * 12 + k] - gmin[k]) / cs);
            ci[(size_t)t * 3 + k] = c;
            dim[k] = std::max(dim[k], c + 2);
        }
    for (int t = 0; t < nT; t++)
        cell[t] = ci[(size_t)t * 3] + (long long)dim[0] * (ci[(size_t)t * 3 + 1] + (long long)dim[1] * ci[(size_t)t * 3 + 2]);
    std::vector<int> order((size_t)nT);
    for (int t = 0; t < nT; t++) order[t] = t;
    std::sort(order.begin(), order.end(), [&](int a, int b) { return cell[a] != cell[b] ? cell[a] < cell[b] : a < b; });
    std::vector<long long> keys((size_t)nT);
    for (int i = 0; i < nT; i++) keys[i] = cell[order[i]];
    long long count = 0;
    std::vector<int> cand;
    for (int a = 0; a < nT; a++) {
        cand.clear();
        const int *ca = &ci[(size_t)a * 3];
        for (int dz = -1; dz <= 1; dz++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    int cx = ca[0] + dx, cy = ca[1] + dy, cz = ca[2] + dz;
                    if (cx < 0 || cy < 0 || cz < 0) continue;
                    long long key = cx + (long long)dim[0] * (cy + (long long)dim[1] * cz);
                    auto it = std::lower_bound(keys.begin(), keys.end(), key);
                    for (; it != keys.end() && *it == key; ++it) {
                        int b = order[it - keys.begin()];
                        if (b <= a) continue;
                        const double *la = &lo[(size_t)a * 12], *ha = &hi[(size_t)a * 12];
                        const double *lb = &lo[(size_t)b * 12], *hb = &hi[(size_t)b * 12];
                        bool ov = true;
                        for (int d = 0; d < 12 && ov; d++) ov = !(la[d] > hb[d]) && !(ha[d] < lb[d]);
                        if (!ov) continue;
                        if (exclude_shared) {
                            bool sh = false;
                            for (int i = 0; i < 4 && !sh; i++)
                                for (int j = 0; j < 4 && !sh; j++) sh = ie[a + (size_t)es * i] == ie[b + (size_t)es * j];
                            if (sh) continue;
                        }
                        cand.push_back(b);
                    }
                }
        std::sort(cand.begin(), cand.end());
        for (int b : cand) {
            if (count < max_pairs) {
                pairs[2 * count] = a;
                pairs[2 * count + 1] = b;
            }
            count++;
        }
    }
    *nOut = count;
    return count > max_pairs ? ICF_E_INDEX : 0;
}

}  // extern "C"
