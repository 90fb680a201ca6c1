// Host-side model of the calibration path: phase matrices, flags, solution
// slices and the pseudo-inverse of the phase matrices.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <queue>
#include <sstream>

#include "leap_internal.h"

namespace leap {

void hip_check(hipError_t e, const char* what, const char* file, int line) {
    if (e != hipSuccess) {
        std::ostringstream ss;
        ss << what << " failed: " << hipGetErrorString(e) << " (" << file << ":" << line << ")";
        throw Error(e == hipErrorOutOfMemory ? LEAP_ERR_OUT_OF_MEMORY : LEAP_ERR_HIP, ss.str());
    }
}

DeviceBuffer::~DeviceBuffer() {
    if (ptr) (void)hipFree(ptr);
}

void DeviceBuffer::reserve(size_t n) {
    if (n <= bytes && ptr) return;
    if (ptr) {
        LEAP_HIP(hipFree(ptr));
        ptr = nullptr;
        bytes = 0;
    }
    if (n == 0) return;
    LEAP_HIP(hipMalloc(&ptr, n));
    bytes = n;
}

PinnedBuffer::~PinnedBuffer() {
    if (ptr) (void)hipHostFree(ptr);
}

void PinnedBuffer::reserve(size_t n) {
    if (n <= bytes && ptr) return;
    if (ptr) {
        LEAP_HIP(hipHostFree(ptr));
        ptr = nullptr;
        bytes = 0;
    }
    if (n == 0) return;
    LEAP_HIP(hipHostMalloc(&ptr, n, hipHostMallocDefault));
    bytes = n;
}

std::vector<double> PhaseMatrix::dense() const {
    std::vector<double> a(size_t(rows) * cols, 0.0);
    for (int32_t k = 0; k + 1 < rows; ++k) {
        a[size_t(ant1[k]) * rows + k] = 1.0;
        a[size_t(ant2[k]) * rows + k] = -1.0;
    }
    a[size_t(constraint) * rows + (rows - 1)] = 1.0;
    return a;
}

// PhaseMatrixFunction (algorithm/cpu/PhaseMatrixFunction.cc:35-93).  Rows are
// kept sparse (two antenna indices each): every A row holds +1 at antenna1 and
// -1 at antenna2 of one unflagged cross baseline, then one constraint row.
PhaseMatrix phase_matrix(const int32_t* a1, const int32_t* a2, const uint8_t* fg, int32_t n,
                         int32_t ref_ant) {
    if (ref_ant >= 0 && ref_ant >= n) {  // :46-51 (bounds against the baseline count)
        std::ostringstream ss;
        ss << "refAnt " << ref_ant << " is out of bounds";
        throw Error(LEAP_ERR_INVALID_ARGUMENT, ss.str());
    }
    if (ref_ant >= 0 && fg[ref_ant]) {  // :52-57 (flag looked up at index refAnt)
        std::ostringstream ss;
        ss << "refAnt " << ref_ant << " is flagged";
        throw Error(LEAP_ERR_INVALID_ARGUMENT, ss.str());
    }
    require(n > 0, LEAP_ERR_INVALID_ARGUMENT, "a1 and a2 must not be empty");
    PhaseMatrix pm;
    int32_t maxa = 0;
    for (int32_t i = 0; i < n; ++i) {
        require(a1[i] >= 0 && a2[i] >= 0, LEAP_ERR_INVALID_ARGUMENT, "negative antenna index");
        maxa = std::max(maxa, std::max(a1[i], a2[i]));
    }
    pm.cols = maxa + 1;  // :59
    for (int32_t i = 0; i < n; ++i) {  // :65-79
        if (a1[i] == a2[i] || fg[i]) continue;
        if (ref_ant >= 0 && a1[i] != ref_ant && a2[i] != ref_ant) continue;
        pm.I.push_back(i);
        pm.ant1.push_back(a1[i]);
        pm.ant2.push_back(a2[i]);
    }
    pm.constraint = ref_ant < 0 ? 0 : ref_ant;  // :81-86
    pm.rows = int32_t(pm.I.size()) + 1;
    return pm;
}

namespace {

// In-place inverse of a symmetric positive definite n x n matrix (column-major)
// by Cholesky factorisation M = L L^T, then M^-1 = L^-T L^-1.
void spd_inverse(std::vector<double>& M, int32_t n) {
    auto at = [&](int32_t i, int32_t j) -> double& { return M[size_t(j) * n + i]; };
    for (int32_t j = 0; j < n; ++j) {  // L in the lower triangle
        double d = at(j, j);
        for (int32_t k = 0; k < j; ++k) d -= at(j, k) * at(j, k);
        require(d > 0.0, LEAP_ERR_INTERNAL, "phase matrix Gram block is not positive definite");
        const double ljj = std::sqrt(d);
        at(j, j) = ljj;
        for (int32_t i = j + 1; i < n; ++i) {
            double s = at(i, j);
            for (int32_t k = 0; k < j; ++k) s -= at(i, k) * at(j, k);
            at(i, j) = s / ljj;
        }
    }
    // W = L^-1 (lower triangular), by forward substitution column by column.
    std::vector<double> W(size_t(n) * n, 0.0);
    auto w = [&](int32_t i, int32_t j) -> double& { return W[size_t(j) * n + i]; };
    for (int32_t j = 0; j < n; ++j) {
        w(j, j) = 1.0 / at(j, j);
        for (int32_t i = j + 1; i < n; ++i) {
            double s = 0.0;
            for (int32_t k = j; k < i; ++k) s -= at(i, k) * w(k, j);
            w(i, j) = s / at(i, i);
        }
    }
    // M^-1 = W^T W
    for (int32_t j = 0; j < n; ++j) {
        for (int32_t i = j; i < n; ++i) {
            double s = 0.0;
            for (int32_t k = i; k < n; ++k) s += w(k, i) * w(k, j);
            at(i, j) = s;
            at(j, i) = s;
        }
    }
}

struct UnionFind {
    std::vector<int32_t> p;
    explicit UnionFind(int32_t n) : p(static_cast<size_t>(n)) { std::iota(p.begin(), p.end(), 0); }
    int32_t find(int32_t x) {
        while (p[x] != x) x = p[x] = p[p[x]];
        return x;
    }
    void unite(int32_t a, int32_t b) {
        a = find(a);
        b = find(b);
        if (a != b) p[std::max(a, b)] = std::min(a, b);
    }
};

}  // namespace

// Moore-Penrose inverse of G = A^T A, where A holds the rows e_a1 - e_a2 of the
// used baselines plus the constraint row e_c.
//
// G is an integer matrix (graph Laplacian of the baseline graph plus e_c e_c^T),
// formed exactly.  Its null space is spanned by the indicator vectors of the
// connected components of the baseline graph that do not contain c (an unused
// antenna is a component of its own with a zero column).  So, per component S:
//   c in S:       G_SS is positive definite       ->  G+_SS = G_SS^-1
//   c not in S:   G_SS = L_S with null vector 1   ->  G+_SS = (L_S + J/|S|)^-1 - J/|S|
//   isolated:     G_SS = 0                        ->  G+_SS = 0
// and G+ is zero between components.  Then pinv(A) = G+ A^T.  This is the
// matrix the reference's SVD pseudo-inverse computes (math/cpu/matrix_invert.h:49-56):
// the nonzero singular values of an incidence matrix are bounded well away
// from the reference's cut max(m, n) * eps * sigma_max, so the two agree on
// the rank, and the rank is found here combinatorially instead of by a cut.
std::vector<double> gram_pinv(const PhaseMatrix& pm) {
    const int32_t N = pm.cols;
    std::vector<double> G(size_t(N) * N, 0.0);
    auto g = [&](int32_t i, int32_t j) -> double& { return G[size_t(j) * N + i]; };
    UnionFind uf(N);
    std::vector<int32_t> degree(static_cast<size_t>(N), 0);
    for (size_t k = 0; k < pm.ant1.size(); ++k) {
        const int32_t a = pm.ant1[k], b = pm.ant2[k];
        g(a, a) += 1.0;
        g(b, b) += 1.0;
        g(a, b) -= 1.0;
        g(b, a) -= 1.0;
        ++degree[a];
        ++degree[b];
        uf.unite(a, b);
    }
    g(pm.constraint, pm.constraint) += 1.0;
    std::vector<double> Gp(size_t(N) * N, 0.0);
    std::vector<std::vector<int32_t>> comps(static_cast<size_t>(N));
    for (int32_t a = 0; a < N; ++a) comps[uf.find(a)].push_back(a);
    for (const auto& S : comps) {
        const int32_t n = int32_t(S.size());
        if (n == 0) continue;
        const bool has_c = std::find(S.begin(), S.end(), pm.constraint) != S.end();
        if (n == 1 && !has_c && degree[S[0]] == 0) continue;  // unused antenna
        std::vector<double> M(size_t(n) * n);
        for (int32_t j = 0; j < n; ++j)
            for (int32_t i = 0; i < n; ++i) M[size_t(j) * n + i] = g(S[i], S[j]);
        const double inv_n = 1.0 / double(n);
        if (!has_c)
            for (auto& x : M) x += inv_n;
        spd_inverse(M, n);
        if (!has_c)
            for (auto& x : M) x -= inv_n;
        for (int32_t j = 0; j < n; ++j)
            for (int32_t i = 0; i < n; ++i) Gp[size_t(S[j]) * N + S[i]] = M[size_t(j) * n + i];
    }
    return Gp;
}

GramComponents gram_components(const PhaseMatrix& pm) {
    const int32_t N = pm.cols;
    UnionFind uf(N);
    std::vector<int32_t> degree(static_cast<size_t>(N), 0);
    for (size_t k = 0; k < pm.ant1.size(); ++k) {
        uf.unite(pm.ant1[k], pm.ant2[k]);
        ++degree[pm.ant1[k]];
        ++degree[pm.ant2[k]];
    }
    std::vector<std::vector<int32_t>> comps(static_cast<size_t>(N));
    for (int32_t a = 0; a < N; ++a) comps[uf.find(a)].push_back(a);
    GramComponents out;
    out.off.push_back(0);
    for (const auto& S : comps) {
        if (S.empty()) continue;
        const bool has_c = std::find(S.begin(), S.end(), pm.constraint) != S.end();
        if (S.size() == 1 && !has_c && degree[S[0]] == 0) continue;  // unused antenna: G+ = 0
        out.members.insert(out.members.end(), S.begin(), S.end());
        out.off.push_back(int32_t(out.members.size()));
        out.has_c.push_back(has_c ? 1 : 0);
    }
    return out;
}

// Components of the baseline graph that do not hold the constraint antenna
// (an unused antenna is one of its own): the dimension of the null space of A.
// Ad A is the projection I - sum_S 1_S 1_S^T / |S| over them, so
// ||Ad A - I||_F = sqrt(k) and ||Ad A||_F = sqrt(N - k): the reference's test
// (Ad A).isApprox(I, 0.001) (model/cpu/MetaData.cc:129-136) fails exactly when
// k >= 1 (for N < 1e6 k).
int32_t null_components(const PhaseMatrix& pm) {
    const int32_t N = pm.cols;
    UnionFind uf(N);
    for (size_t k = 0; k < pm.ant1.size(); ++k) uf.unite(pm.ant1[k], pm.ant2[k]);
    const int32_t c = uf.find(pm.constraint);
    int32_t k = 0;
    for (int32_t a = 0; a < N; ++a)
        if (uf.find(a) == a && a != c) ++k;
    return k;
}

// MeasurementSet::GetFilteredBaselines (ms/MeasurementSet.cc:181-240).
std::vector<uint8_t> filtered_baselines(const leap_ms_view& ms, double threshold) {
    const uint32_t nbl = ms.num_baselines;
    std::vector<uint8_t> out(nbl, 0);
    for (uint32_t i = 0; i < nbl; ++i) out[i] = ms.flag[i] ? 1 : 0;  // FLAG[row][0][0], rows [0, nbl)
    if (threshold > 0.0) {
        // GetShortBaselines (:205-228) reads UVW through a slicer that keeps u
        // only, into a (1 x rows) matrix it then indexes as uv(i, 0), uv(i, 1):
        // in casacore's unchecked release indexing those are u of rows i and i+1.
        // Kept as such; the read past the last row (one timestep only) is 0.
        for (uint32_t i = 0; i < nbl; ++i) {
            const double x = ms.uvw[size_t(i) * 3];
            const double y = (i + 1 < ms.num_rows) ? ms.uvw[size_t(i + 1) * 3] : 0.0;
            if (std::sqrt(x * x + y * y) < threshold) out[i] = 1;
        }
    }
    return out;
}

// Slice ctor (common/Slice.cc:36-63), Slice::Evaluate (Slice.h:47-54), Range ctor (Range.h:40-49).
void evaluate_slice(int32_t start, int32_t interval, int32_t end, int32_t T, int32_t out[3]) {
    require(start >= 0, LEAP_ERR_RANGE, "expected a positive integer");
    require(interval >= -1, LEAP_ERR_RANGE, "expected a positive integer or -1");
    require(interval != 0, LEAP_ERR_RANGE, "expected a non zero integer");
    require(end >= -1, LEAP_ERR_RANGE, "expected a positive integer or -1");
    require(end == -1 || start < end, LEAP_ERR_RANGE, "range start must be greater than end");
    if (end != -1) {
        if (interval == -1) interval = end - start;
        require(interval <= end - start, LEAP_ERR_RANGE, "range increment out of bounds");
    }
    out[0] = start == -1 ? T : start;
    out[1] = interval == -1 ? T : interval;
    out[2] = end == -1 ? T : end;
    require(out[0] >= 0 && out[1] >= 1 && out[2] >= 0, LEAP_ERR_RANGE, "expected a positive integer");
}

// MeasurementSet::Validate (ms/MeasurementSet.cc:64-103), the checks that apply to a view.
void validate_view(const leap_ms_view& ms) {
    require(ms.antenna1 && ms.antenna2 && ms.flag && ms.uvw && ms.time, LEAP_ERR_INVALID_ARGUMENT,
            "measurement set view has null columns");
    require(ms.num_rows >= ms.num_baselines, LEAP_ERR_RANGE,
            "invalid number of rows, expected >=" + std::to_string(ms.num_baselines) + ", got " +
                std::to_string(ms.num_rows));
    require(ms.num_baselines != 0 && ms.num_rows % ms.num_baselines == 0, LEAP_ERR_RANGE,
            "number of rows not an integer multiple of baselines");
    require(ms.num_channels >= 1, LEAP_ERR_RANGE, "measurement set has no channels");
    require(ms.num_pols >= 1 && ms.num_pols <= 4, LEAP_ERR_NOT_IMPLEMENTED,
            "polarisation counts above 4 are not supported");
    require(ms.vis_dtype == LEAP_VIS_C128 || ms.vis_dtype == LEAP_VIS_C64, LEAP_ERR_INVALID_ARGUMENT,
            "unknown visibility dtype");
}

uint64_t hash_phase_matrix(const PhaseMatrix& pm) {
    uint64_t h = 1469598103934665603ull;
    auto add = [&](uint64_t v) {
        for (int i = 0; i < 8; ++i) {
            h ^= (v >> (8 * i)) & 0xff;
            h *= 1099511628211ull;
        }
    };
    add(uint64_t(pm.rows));
    add(uint64_t(pm.cols));
    add(uint64_t(pm.constraint));
    for (size_t k = 0; k < pm.ant1.size(); ++k) {
        add(uint64_t(uint32_t(pm.ant1[k])) | (uint64_t(uint32_t(pm.ant2[k])) << 32));
    }
    return h;
}

CapturedSequence::~CapturedSequence() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
}

bool multi_tile_supported(int32_t tile) {
    return tile == 321 || tile == 641 || tile == 1281 || tile == 1282 || tile == 2562;
}

// Work of one (baseline, direction block) item of the multi-set kernel: per
// group of RP = 256 / (DB / ND) rows, ceil(C / 2) channel pairs on the FMA
// chains plus the staging of each 256-channel chunk and the row set-up.
double multi_item_cost(uint32_t T, uint32_t C, int32_t tile) {
    require(multi_tile_supported(tile), LEAP_ERR_INVALID_ARGUMENT,
            "unsupported multi-set tile " + std::to_string(tile));
    const uint32_t lanes = uint32_t(tile / 10) / uint32_t(tile % 10);
    const uint32_t rp = 256 / lanes;
    const double groups = double((T + rp - 1) / rp);
    const double chunks = double((C + 255) / 256);
    return groups * (double((C + 1) / 2) + 48.0 * chunks + 16.0) + 32.0;
}

// Longest processing time first: items in decreasing cost, each to the
// least-loaded workgroup (ties to the lower index), so the plan is a
// deterministic function of the shapes.
MultiPlan plan_multi(const std::vector<MultiShape>& sets, int32_t tile, uint32_t n_blocks) {
    require(n_blocks >= 1, LEAP_ERR_INVALID_ARGUMENT, "the schedule needs at least one workgroup");
    const uint32_t db = uint32_t(tile / 10);
    std::vector<uint32_t> order(sets.size());
    std::vector<double> cost(sets.size(), 0.0);
    size_t n_items = 0;
    for (size_t s = 0; s < sets.size(); ++s) {
        order[s] = uint32_t(s);
        cost[s] = multi_item_cost(sets[s].T, sets[s].C, tile);
        if (sets[s].D > 0) {
            require(sets[s].nbl > 0, LEAP_ERR_INVALID_ARGUMENT, "a set has no baselines");
            n_items += size_t(sets[s].nbl) * ((sets[s].D + db - 1) / db);
        }
    }
    require(n_items < (size_t(1) << 32), LEAP_ERR_RANGE, "too many work items for one launch");
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return cost[a] > cost[b]; });
    using Slot = std::pair<double, uint32_t>;  // (load, workgroup)
    std::priority_queue<Slot, std::vector<Slot>, std::greater<Slot>> heap;
    for (uint32_t g = 0; g < n_blocks; ++g) heap.emplace(0.0, g);
    std::vector<uint32_t> owner(n_items), count(n_blocks, 0);
    std::vector<uint32_t> iset(n_items), ilocal(n_items);
    size_t k = 0;
    for (uint32_t s : order) {
        if (sets[s].D == 0) continue;
        const uint32_t n = sets[s].nbl * ((sets[s].D + db - 1) / db);
        for (uint32_t i = 0; i < n; ++i, ++k) {
            Slot top = heap.top();
            heap.pop();
            owner[k] = top.second;
            ++count[top.second];
            iset[k] = s;
            ilocal[k] = i;
            top.first += cost[s];
            heap.push(top);
        }
    }
    MultiPlan plan;
    plan.makespan = 0.0;
    while (!heap.empty()) {
        plan.makespan = std::max(plan.makespan, heap.top().first);
        heap.pop();
    }
    plan.off.assign(size_t(n_blocks) + 1, 0);
    for (uint32_t g = 0; g < n_blocks; ++g) plan.off[g + 1] = plan.off[g] + count[g];
    plan.items.resize(n_items);
    std::vector<uint32_t> fill(plan.off.begin(), plan.off.end() - 1);
    for (size_t i = 0; i < n_items; ++i) plan.items[fill[owner[i]]++] = {iset[i], ilocal[i]};
    return plan;
}

}  // namespace leap
