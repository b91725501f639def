// fusion.cpp -- k-qubit gate fusion for the MI355X engine.
//
// The reference fuses only runs of single-qubit gates on one qubit into a
// 2x2 and treats every 2-qubit gate as a barrier (src/optimization/fusion/
// fusion.c:247-292).  Here any run of gates whose qubits fit in k <= 5 is
// folded into one 2^k x 2^k unitary composed in FP64 on the host, then
// applied in ONE pass over the state, so a block of g gates costs one HBM
// read + write instead of g.
#include "fusion.h"

#include <math.h>
#include <string.h>

#include <algorithm>

#include "sv_internal.h"

namespace mlx {

hipError_t launch_dense(const SvDev &s, int k, const int *qubits, const double *mat);
hipError_t launch_dense_dev(const SvDev &s, int k, const int *qubits, const void *d_mat);

namespace {

// What a gate does on its qubits: a 2x2 on `target` where every control is 1,
// or an exchange of two targets where every control is 1.
struct GateAction {
    int  nctrl;
    int  ctrl[2];
    int  target;
    bool is_swap;
    int  target2;
    cd   u[4];
};

GateAction action_of(const moonlab_hip_gate_t &g)
{
    GateAction a{};
    a.target = g.target;
    const cd I(0.0, 1.0);
    auto set = [&](cd u00, cd u01, cd u10, cd u11) {
        a.u[0] = u00; a.u[1] = u01; a.u[2] = u10; a.u[3] = u11;
    };
    const double c = cos(g.theta / 2.0), s = sin(g.theta / 2.0);
    switch (g.type) {
    case MOONLAB_HIP_GATE_I: set(1, 0, 0, 1); break;
    case MOONLAB_HIP_GATE_X: set(0, 1, 1, 0); break;
    case MOONLAB_HIP_GATE_Y: set(0, -I, I, 0); break;
    case MOONLAB_HIP_GATE_Z: set(1, 0, 0, -1); break;
    case MOONLAB_HIP_GATE_H: {
        const double r = 0.70710678118654752440;
        set(r, r, r, -r);
        break;
    }
    case MOONLAB_HIP_GATE_S: set(1, 0, 0, I); break;
    case MOONLAB_HIP_GATE_SDG: set(1, 0, 0, -I); break;
    case MOONLAB_HIP_GATE_T: set(1, 0, 0, cd(cos(M_PI / 4.0), sin(M_PI / 4.0))); break;
    case MOONLAB_HIP_GATE_TDG: set(1, 0, 0, cd(cos(-M_PI / 4.0), sin(-M_PI / 4.0))); break;
    case MOONLAB_HIP_GATE_PHASE: set(1, 0, 0, cd(cos(g.theta), sin(g.theta))); break;
    case MOONLAB_HIP_GATE_RX: set(c, cd(0, -s), cd(0, -s), c); break;
    case MOONLAB_HIP_GATE_RY: set(c, -s, s, c); break;
    case MOONLAB_HIP_GATE_RZ: set(cd(c, -s), 0, 0, cd(c, s)); break;
    case MOONLAB_HIP_GATE_U3:
        set(c, -cd(cos(g.lam), sin(g.lam)) * s, cd(cos(g.phi), sin(g.phi)) * s,
            cd(cos(g.phi + g.lam), sin(g.phi + g.lam)) * c);
        break;
    default: break;
    }
    switch (g.type) {
    case MOONLAB_HIP_GATE_CNOT: a.nctrl = 1; a.ctrl[0] = g.control; set(0, 1, 1, 0); break;
    case MOONLAB_HIP_GATE_CY: a.nctrl = 1; a.ctrl[0] = g.control; set(0, I, -I, 0); break;
    case MOONLAB_HIP_GATE_CY_TRUE: a.nctrl = 1; a.ctrl[0] = g.control; set(0, -I, I, 0); break;
    case MOONLAB_HIP_GATE_CZ: a.nctrl = 1; a.ctrl[0] = g.control; set(1, 0, 0, -1); break;
    case MOONLAB_HIP_GATE_CPHASE:
        a.nctrl = 1; a.ctrl[0] = g.control; set(1, 0, 0, cd(cos(g.theta), sin(g.theta)));
        break;
    case MOONLAB_HIP_GATE_CRX: a.nctrl = 1; a.ctrl[0] = g.control; set(c, cd(0, -s), cd(0, -s), c); break;
    case MOONLAB_HIP_GATE_CRY: a.nctrl = 1; a.ctrl[0] = g.control; set(c, -s, s, c); break;
    case MOONLAB_HIP_GATE_CRZ: a.nctrl = 1; a.ctrl[0] = g.control; set(cd(c, -s), 0, 0, cd(c, s)); break;
    case MOONLAB_HIP_GATE_TOFFOLI:
        a.nctrl = 2; a.ctrl[0] = g.control; a.ctrl[1] = g.extra; set(0, 1, 1, 0);
        break;
    case MOONLAB_HIP_GATE_SWAP: a.is_swap = true; a.target = g.control; a.target2 = g.target; break;
    case MOONLAB_HIP_GATE_FREDKIN:
        a.nctrl = 1; a.ctrl[0] = g.control; a.is_swap = true; a.target2 = g.extra;
        break;
    default: break;
    }
    return a;
}

// Apply an action to a 2^k vector whose local bit j is global qubit qs[j].
void apply_local(const GateAction &a, const int *qs, int k, cd *v)
{
    auto local = [&](int q) {
        for (int j = 0; j < k; j++)
            if (qs[j] == q) return j;
        return -1;
    };
    uint64_t cm = 0;
    for (int i = 0; i < a.nctrl; i++) cm |= 1ull << local(a.ctrl[i]);
    const int dim = 1 << k;
    if (a.is_swap) {
        const uint64_t b1 = 1ull << local(a.target), b2 = 1ull << local(a.target2);
        for (int i = 0; i < dim; i++)
            if ((i & cm) == cm && (i & b1) && !(i & b2)) std::swap(v[i], v[(i ^ b1) | b2]);
        return;
    }
    const uint64_t tb = 1ull << local(a.target);
    for (int i = 0; i < dim; i++) {
        if ((i & cm) != cm || (i & tb)) continue;
        const cd x0 = v[i], x1 = v[i | tb];
        v[i] = a.u[0] * x0 + a.u[1] * x1;
        v[i | tb] = a.u[2] * x0 + a.u[3] * x1;
    }
}

}  // namespace

bool gate_matrix_1q(const moonlab_hip_gate_t &g, cd u[4])
{
    int q[3];
    if (gate_qubits(g, q) != 1) return false;
    const GateAction a = action_of(g);
    for (int i = 0; i < 4; i++) u[i] = a.u[i];
    return true;
}

int gate_qubits(const moonlab_hip_gate_t &g, int q[3])
{
    int n = 0;
    q[n++] = g.target;
    const bool two = (g.type >= MOONLAB_HIP_GATE_CNOT && g.type <= MOONLAB_HIP_GATE_SWAP) ||
                     g.type >= MOONLAB_HIP_GATE_CPHASE;
    if (two) q[n++] = g.control;
    if (g.type == MOONLAB_HIP_GATE_TOFFOLI || g.type == MOONLAB_HIP_GATE_FREDKIN) q[n++] = g.extra;
    std::sort(q, q + n);
    return n;
}

bool gate_valid(const moonlab_hip_gate_t &g, int n)
{
    if (g.type < 0 || g.type > MOONLAB_HIP_GATE_FREDKIN) return false;
    int q[3];
    const int k = gate_qubits(g, q);
    for (int j = 0; j < k; j++)
        if (q[j] < 0 || q[j] >= n || (j && q[j] == q[j - 1])) return false;
    return true;
}

void gate_dense(const moonlab_hip_gate_t &g, cd *m)
{
    int q[3];
    const int k = gate_qubits(g, q);
    const GateAction a = action_of(g);
    const int dim = 1 << k;
    std::vector<cd> col((size_t)dim);
    for (int c = 0; c < dim; c++) {
        std::fill(col.begin(), col.end(), cd(0.0));
        col[(size_t)c] = 1.0;
        apply_local(a, q, k, col.data());
        for (int r = 0; r < dim; r++) m[(size_t)r * dim + c] = col[(size_t)r];
    }
}

namespace {

// Compose a block's unitary in FP64: the identity, then each member gate
// left-multiplied in circuit order.
void compose_block(FusedBlock &b, const std::vector<const moonlab_hip_gate_t *> &members)
{
    const int k = (int)b.qubits.size(), dim = 1 << k;
    b.mat.assign((size_t)dim * dim, cd(0.0));
    for (int i = 0; i < dim; i++) b.mat[(size_t)i * dim + i] = 1.0;
    std::vector<cd> col((size_t)dim);
    for (const moonlab_hip_gate_t *gp : members) {
        if (gp->type == MOONLAB_HIP_GATE_I) continue;
        const GateAction a = action_of(*gp);
        for (int c = 0; c < dim; c++) {
            for (int r = 0; r < dim; r++) col[(size_t)r] = b.mat[(size_t)r * dim + c];
            apply_local(a, b.qubits.data(), k, col.data());
            for (int r = 0; r < dim; r++) b.mat[(size_t)r * dim + c] = col[(size_t)r];
        }
    }
    b.num_gates = (int)members.size();
    b.only_gate = members.size() == 1 ? members.front() : nullptr;
}

}  // namespace

// Blocks on disjoint qubits commute, so several blocks stay open at once, one
// per disjoint qubit set: a gate joins (and merges) the open blocks it
// touches when their union still fits kmax; otherwise the largest of those
// blocks are emitted until it fits.  A block is emitted only once nothing
// later in the circuit can join it, and every gate that shares a qubit with
// it and comes later lands in a block emitted after it, so the emitted order
// equals the circuit up to commuting disjoint blocks.
std::vector<FusedBlock> plan_fusion(const moonlab_hip_gate_t *g, int ng, int kmax)
{
    struct Open {
        uint64_t                                mask = 0;
        std::vector<const moonlab_hip_gate_t *> members;
        bool                                    live = false;
    };
    std::vector<Open> open;
    int owner[64];
    std::fill(owner, owner + 64, -1);
    std::vector<FusedBlock> blocks;
    auto emit = [&](int id) {
        Open &o = open[(size_t)id];
        FusedBlock b;
        for (int q = 0; q < 64; q++)
            if (o.mask & (1ull << q)) {
                b.qubits.push_back(q);
                owner[q] = -1;
            }
        compose_block(b, o.members);
        blocks.push_back(std::move(b));
        o.live = false;
        o.members.clear();
        o.mask = 0;
    };
    for (int i = 0; i < ng; i++) {
        int q[3];
        const int nq = gate_qubits(g[i], q);
        uint64_t sm = 0;
        for (int j = 0; j < nq; j++) sm |= 1ull << q[j];
        std::vector<int> touched;
        for (int j = 0; j < nq; j++) {
            const int o = owner[q[j]];
            if (o >= 0 && std::find(touched.begin(), touched.end(), o) == touched.end()) touched.push_back(o);
        }
        auto union_mask = [&]() {
            uint64_t u = sm;
            for (int o : touched) u |= open[(size_t)o].mask;
            return u;
        };
        uint64_t u = union_mask();
        if (__builtin_popcountll(u) > kmax) {
            std::sort(touched.begin(), touched.end(), [&](int x, int y) {
                const int px = __builtin_popcountll(open[(size_t)x].mask), py = __builtin_popcountll(open[(size_t)y].mask);
                return px != py ? px > py : x < y;
            });
            // a gate wider than kmax (Toffoli at kmax 2) forms a block of its own
            while (__builtin_popcountll(u) > kmax && !touched.empty()) {
                emit(touched.front());
                touched.erase(touched.begin());
                u = union_mask();
            }
        }
        // merge what is left into a new open block, members in circuit order
        Open nb;
        nb.live = true;
        nb.mask = u;
        std::sort(touched.begin(), touched.end());   // creation order
        for (int o : touched) {
            Open &src = open[(size_t)o];
            nb.members.insert(nb.members.end(), src.members.begin(), src.members.end());
            src.live = false;
            src.members.clear();
            src.mask = 0;
        }
        std::sort(nb.members.begin(), nb.members.end());   // pointers into g: circuit order
        nb.members.push_back(&g[i]);
        open.push_back(std::move(nb));
        const int id = (int)open.size() - 1;
        for (int qq = 0; qq < 64; qq++)
            if (u & (1ull << qq)) owner[qq] = id;
    }
    for (size_t id = 0; id < open.size(); id++)
        if (open[id].live) emit((int)id);
    return blocks;
}

// Window planning: gates commute when they share no qubit, so a block takes,
// in circuit order, every pending gate whose qubits fit its qubit set and
// that shares no qubit with a gate already passed over; the passed-over
// gates, still in order, are the rest of the circuit.  The set is chosen by
// lookahead: besides the set grown from the first pending gate, the first
// `seeds` pending gates seed windows of kmax neighbouring qubits (starting
// at, centred on and ending at the gate's lowest qubit), and the set taking
// the most gates wins.  On nearest-neighbour circuits a window runs several
// layers inside its light cone, where the open-block planner closes a block
// per layer.
std::vector<FusedBlock> plan_fusion_windows(const moonlab_hip_gate_t *g, int ng, int kmax, int seeds, int n)
{
    std::vector<FusedBlock> blocks;
    std::vector<const moonlab_hip_gate_t *> rest, best_take, best_defer;
    std::vector<uint64_t> qm((size_t)ng);
    for (int i = 0; i < ng; i++) {
        int q[3];
        const int nq = gate_qubits(g[i], q);
        uint64_t m = 0;
        for (int j = 0; j < nq; j++) m |= 1ull << q[j];
        qm[(size_t)i] = m;
        rest.push_back(&g[i]);
    }
    auto mask = [&](const moonlab_hip_gate_t *x) { return qm[(size_t)(x - g)]; };
    // gates a qubit set starting from S0 takes among rest[0, lim) (growing
    // while it has room when grow is set); with take / defer, the full split
    auto scan = [&](uint64_t S0, bool grow, size_t lim, std::vector<const moonlab_hip_gate_t *> *tk,
                    std::vector<const moonlab_hip_gate_t *> *df) {
        uint64_t S = S0, blocked = 0;
        size_t taken = 0;
        for (size_t i = 0; i < lim; i++) {
            const moonlab_hip_gate_t *x = rest[i];
            const uint64_t m = mask(x);
            const bool fits = grow ? __builtin_popcountll(S | m) <= kmax : (m & ~S) == 0;
            if (!(m & blocked) && fits) {
                S |= m;
                taken++;
                if (tk) tk->push_back(x);
            } else {
                blocked |= m;
                if (df) df->push_back(x);
            }
        }
        return taken;
    };
    // candidates are compared on the first 4096 pending gates, which bounds
    // planning time on very long circuits
    constexpr size_t kLookaheadGates = 4096;
    std::vector<uint64_t> tried;
    while (!rest.empty()) {
        const size_t lim = rest.size() < kLookaheadGates ? rest.size() : kLookaheadGates;
        size_t best = scan(0, true, lim, nullptr, nullptr);
        uint64_t bestS0 = 0;
        bool best_grow = true;
        tried.clear();
        int used = 0;
        for (size_t i = 0; i < rest.size() && used < seeds; i++) {
            const uint64_t m = mask(rest[i]);
            if (!m || __builtin_popcountll(m) > kmax) continue;
            used++;
            const int q = __builtin_ctzll(m);
            for (const int lo0 : {q, q - kmax / 2, q - kmax + 1}) {
                uint64_t S0 = m;
                const int lo = lo0 < 0 ? 0 : lo0;
                for (int k = lo; k < n && __builtin_popcountll(S0) < kmax; k++) S0 |= 1ull << k;
                for (int k = lo - 1; k >= 0 && __builtin_popcountll(S0) < kmax; k--) S0 |= 1ull << k;
                if (std::find(tried.begin(), tried.end(), S0) != tried.end()) continue;
                tried.push_back(S0);
                const size_t c = scan(S0, false, lim, nullptr, nullptr);
                if (c > best) {
                    best = c;
                    bestS0 = S0;
                    best_grow = false;
                }
            }
        }
        best_take.clear();
        best_defer.clear();
        scan(bestS0, best_grow, rest.size(), &best_take, &best_defer);
        if (best_take.empty()) {   // a gate wider than kmax stands alone
            best_take.assign(rest.begin(), rest.begin() + 1);
            best_defer.assign(rest.begin() + 1, rest.end());
        }
        FusedBlock b;
        uint64_t u = 0;
        for (const moonlab_hip_gate_t *x : best_take) u |= mask(x);
        for (int q = 0; q < 64; q++)
            if (u & (1ull << q)) b.qubits.push_back(q);
        compose_block(b, best_take);
        blocks.push_back(std::move(b));
        rest.swap(best_defer);
    }
    return blocks;
}

// The open-block plan or the window plan, whichever has fewer blocks (each
// block is one pass over the state).
std::vector<FusedBlock> plan_fusion_best(const moonlab_hip_gate_t *g, int ng, int kmax)
{
    std::vector<FusedBlock> a = plan_fusion(g, ng, kmax);
    int n = 0;
    for (int i = 0; i < ng; i++) {
        int q[3];
        const int nq = gate_qubits(g[i], q);
        for (int j = 0; j < nq; j++) n = std::max(n, q[j] + 1);
    }
    std::vector<FusedBlock> b = plan_fusion_windows(g, ng, kmax, kFusionSeeds, n);
    return b.size() < a.size() ? b : a;
}

int apply_fused(moonlab_hip_state_t *s, const moonlab_hip_gate_t *g, int ng, int kmax)
{
    for (int i = 0; i < ng; i++) {
        int q[3];
        const int nq = gate_qubits(g[i], q);
        for (int j = 0; j < nq; j++)
            if (q[j] < 0 || q[j] >= (int)moonlab_hip_state_n_qubits(s)) return MOONLAB_HIP_ERR_INVALID_QUBIT;
        for (int j = 1; j < nq; j++)
            if (q[j] == q[j - 1]) return MOONLAB_HIP_ERR_INVALID_QUBIT;
        if (g[i].type < 0 || g[i].type > MOONLAB_HIP_GATE_FREDKIN) return MOONLAB_HIP_ERR_BAD_ARG;
    }
    if (kmax < 1) kmax = 1;
    if (kmax > 5) kmax = 5;
    const std::vector<FusedBlock> blocks = plan_fusion_best(g, ng, kmax);
    const SvDev d{moonlab_hip_state_device_ptr(s), (int)moonlab_hip_state_n_qubits(s), moonlab_hip_state_precision(s),
                  (hipStream_t)moonlab_hip_state_stream(s)};
    const bool f = d.prec == MOONLAB_HIP_C64;
    const size_t ab = f ? 8 : 16;
    // every matrix of 3+ qubits goes to the device in one copy
    std::vector<size_t> off(blocks.size(), 0);
    size_t total = 0;
    for (size_t i = 0; i < blocks.size(); i++)
        if (blocks[i].qubits.size() >= 3 && !blocks[i].only_gate) {
            off[i] = total;
            total += blocks[i].mat.size() * ab;
        }
    std::vector<unsigned char> host(total);
    for (size_t i = 0; i < blocks.size(); i++) {
        if (blocks[i].qubits.size() < 3 || blocks[i].only_gate) continue;
        const std::vector<cd> &m = blocks[i].mat;
        if (f) {
            float *p = (float *)(host.data() + off[i]);
            for (size_t j = 0; j < m.size(); j++) { p[2 * j] = (float)m[j].real(); p[2 * j + 1] = (float)m[j].imag(); }
        } else {
            double *p = (double *)(host.data() + off[i]);
            for (size_t j = 0; j < m.size(); j++) { p[2 * j] = m[j].real(); p[2 * j + 1] = m[j].imag(); }
        }
    }
    void *dm = nullptr;
    hipError_t e = hipSuccess;
    if (total) {
        e = hipMallocAsync(&dm, total, d.stream);
        if (e == hipSuccess) e = hipMemcpyAsync(dm, host.data(), total, hipMemcpyHostToDevice, d.stream);
        if (e == hipSuccess) e = hipStreamSynchronize(d.stream);   // `host` is local
    }
    std::vector<double> m2;
    for (size_t i = 0; e == hipSuccess && i < blocks.size(); i++) {
        const FusedBlock &b = blocks[i];
        const int k = (int)b.qubits.size();
        if (b.only_gate) {   // a lone gate keeps its own kernel (a CZ touches a quarter of the state)
            e = launch_gate(d, *b.only_gate);
            continue;
        }
        if (k >= 3) {
            e = launch_dense_dev(d, k, b.qubits.data(), (const char *)dm + off[i]);
            continue;
        }
        m2.resize(b.mat.size() * 2);
        for (size_t j = 0; j < b.mat.size(); j++) {
            m2[2 * j] = b.mat[j].real();
            m2[2 * j + 1] = b.mat[j].imag();
        }
        e = k == 1 ? launch_matrix_1q(d, b.qubits[0], m2.data()) : launch_matrix_2q(d, b.qubits[0], b.qubits[1], m2.data());
    }
    if (dm) hipFreeAsync(dm, d.stream);
    if (e == hipErrorOutOfMemory) return MOONLAB_HIP_ERR_OUT_OF_MEMORY;
    return e == hipSuccess ? MOONLAB_HIP_OK : MOONLAB_HIP_ERR_DRIVER;
}

}  // namespace mlx

extern "C" {

moonlab_hip_status_t moonlab_hip_apply_dense(moonlab_hip_state_t *s, int k, const int *qubits, const double *mat)
{
    if (!s || !qubits || !mat || k < 1 || k > 5) return MOONLAB_HIP_ERR_BAD_ARG;
    const int n = (int)moonlab_hip_state_n_qubits(s);
    for (int j = 0; j < k; j++) {
        if (qubits[j] < 0 || qubits[j] >= n) return MOONLAB_HIP_ERR_INVALID_QUBIT;
        if (j && qubits[j] <= qubits[j - 1]) return MOONLAB_HIP_ERR_BAD_ARG;
    }
    const mlx::SvDev d{moonlab_hip_state_device_ptr(s), n, moonlab_hip_state_precision(s),
                       (hipStream_t)moonlab_hip_state_stream(s)};
    hipError_t e;
    if (k == 1) {
        e = mlx::launch_matrix_1q(d, qubits[0], mat);
    } else if (k == 2) {
        e = mlx::launch_matrix_2q(d, qubits[0], qubits[1], mat);
    } else {
        e = mlx::launch_dense(d, k, qubits, mat);
    }
    if (e == hipErrorOutOfMemory) return MOONLAB_HIP_ERR_OUT_OF_MEMORY;
    return e == hipSuccess ? MOONLAB_HIP_OK : MOONLAB_HIP_ERR_DRIVER;
}

int moonlab_hip_fusion_plan(const moonlab_hip_gate_t *g, int ng, int kmax, uint64_t *masks, int *counts, double *mats,
                            int max_blocks)
{
    if ((ng > 0 && !g) || ng < 0 || kmax < 1 || kmax > 5 || max_blocks < 0) return MOONLAB_HIP_ERR_BAD_ARG;
    for (int i = 0; i < ng; i++) {
        if (g[i].type < 0 || g[i].type > MOONLAB_HIP_GATE_FREDKIN) return MOONLAB_HIP_ERR_BAD_ARG;
        int q[3];
        const int nq = mlx::gate_qubits(g[i], q);
        for (int j = 0; j < nq; j++)
            if (q[j] < 0 || q[j] > 63 || (j > 0 && q[j] == q[j - 1])) return MOONLAB_HIP_ERR_INVALID_QUBIT;
    }
    const std::vector<mlx::FusedBlock> blocks = mlx::plan_fusion_best(g, ng, kmax);
    for (int i = 0; i < (int)blocks.size() && i < max_blocks; i++) {
        const mlx::FusedBlock &b = blocks[(size_t)i];
        uint64_t m = 0;
        for (int q : b.qubits) m |= 1ull << q;
        if (masks) masks[i] = m;
        if (counts) counts[i] = b.num_gates;
        if (mats)
            for (size_t j = 0; j < b.mat.size(); j++) {
                mats[2048 * (size_t)i + 2 * j] = b.mat[j].real();
                mats[2048 * (size_t)i + 2 * j + 1] = b.mat[j].imag();
            }
    }
    return (int)blocks.size();
}

moonlab_hip_status_t moonlab_hip_apply_gates_fused(moonlab_hip_state_t *s, const moonlab_hip_gate_t *g, int ng,
                                                   int kmax)
{
    if (!s || (ng > 0 && !g) || ng < 0 || kmax < 1 || kmax > 5) return MOONLAB_HIP_ERR_BAD_ARG;
    return (moonlab_hip_status_t)mlx::apply_fused(s, g, ng, kmax);
}

}  // extern "C"
