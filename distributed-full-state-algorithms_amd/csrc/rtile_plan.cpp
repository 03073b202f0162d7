// rtile_plan.cpp -- slot layouts and remaps for register-resident passes
// (k_rtile, kernels_rtile.hip).
//
// A pass's 12 tile bits live on 12 slots: 6 lane slots (0-5), 3 wave slots
// (6-8) and 3 register slots (9-11).  The planner walks the pass's ops in order
// and keeps a layout (slot -> tile bit).  An op needs some of its qubits on
// particular slots: a one-qubit gate's target and a Pauli string's X/Y qubits
// on any lane or register slot; a dense gate's targets on the MFMA slots
// (lane slots 4 and 5, then register slots 9..).  A layout change costs one
// OP_REMAP item, an LDS round trip of the tile, whatever it moves; so each
// layout is chosen to serve the longest window of ops it can (every op's needs
// at once, one dense-gate target set per window), and the next window starts
// with a remap.  Tile bits 0-2 (qubits 0-2, the 128-byte runs of the HBM
// accesses) sit on lane slots in the first window (the load) and must be on
// lane slots at the end (the store), or the pass ends with one more remap.
#include "engine.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace fsa {

namespace {

constexpr int NS = 12;                 // slots (= tile bits)
constexpr int REG0 = 9;                // first register slot
constexpr int MF_SLOTS[5] = {4, 5, 9, 10, 11};

bool is_local(int s) { return s < 6 || s >= REG0; }

struct Layout {
    int bitOf[NS];
    int slotOf[NS];
    Layout() {
        for (int i = 0; i < NS; i++) { bitOf[i] = -1; slotOf[i] = -1; }
    }
    void place(int b, int s) { bitOf[s] = b; slotOf[b] = s; }
    void swap_slots(int s1, int s2) {
        std::swap(bitOf[s1], bitOf[s2]);
        if (bitOf[s1] >= 0) slotOf[bitOf[s1]] = s1;
        if (bitOf[s2] >= 0) slotOf[bitOf[s2]] = s2;
    }
    uint64_t pack() const {
        uint64_t m = 0;
        for (int s = 0; s < NS; s++) m |= (uint64_t)bitOf[s] << (4 * s);
        return m;
    }
    uint32_t slot_mask(uint32_t tileBits) const {
        uint32_t m = 0;
        for (; tileBits; tileBits &= tileBits - 1) m |= 1u << slotOf[__builtin_ctz(tileBits)];
        return m;
    }
};

struct Need {
    bool ok = true;         // the register kernel runs this op
    uint32_t local = 0;     // tile bits that must sit on lane or register slots
    int k = 0;              // dense gate on MFMA: its k target bits in tb
    int tb[5] = {0, 0, 0, 0, 0};
};

uint32_t tile_bits_of(uint64_t gm, const int* pos) {
    uint32_t m = 0;
    for (; gm; gm &= gm - 1) {
        const int p = pos[__builtin_ctzll(gm)];
        if (p >= 0) m |= 1u << p;
    }
    return m;
}

Need need_of(const HOp& op, const int* pos) {
    Need n;
    switch (op.kind) {
    case OP_MAT1: n.local = 1u << pos[op.q[0]]; break;
    case OP_PAULI: n.local = tile_bits_of(op.mxy, pos); break;
    case OP_PHASE: break;
    case OP_MATK:
        if (op.q.size() < 2 || op.q.size() > 5) { n.ok = false; break; }
        n.k = (int)op.q.size();
        for (int i = 0; i < n.k; i++) n.tb[i] = pos[op.q[i]];
        break;
    default: n.ok = false; break;
    }
    return n;
}

// A layout that meets every need of a window of ops at once, or false.
// lanes3: tile bits 0-2 must sit on lane slots (the HBM accesses of the load or
// the store).  keep: the layout in force, whose placements are reused where
// they fit (no cost either way; it keeps layouts stable).
bool window_layout(const std::vector<Need>& needs, size_t i, size_t j, bool lanes3, const Layout* keep,
                   Layout& out) {
    int k = 0, mf[5] = {0, 0, 0, 0, 0};
    uint32_t mfSet = 0, local = 0;
    for (size_t x = i; x <= j; x++) {
        const Need& n = needs[x];
        local |= n.local;
        if (!n.k) continue;
        uint32_t set = 0;
        for (int t = 0; t < n.k; t++) set |= 1u << n.tb[t];
        if (k == 0) {
            k = n.k;
            mfSet = set;
            for (int t = 0; t < k; t++) mf[t] = n.tb[t];
        } else if (set != mfSet) {
            return false;            // two dense gates on different qubits need different slots
        }
    }
    local &= ~mfSet;
    uint32_t lane = lanes3 ? (7u & ~mfSet) : 0u;   // bits that must sit on lane slots
    local &= ~lane;
    Layout L;
    bool slotUsed[NS] = {};
    // dense-gate targets: bits 0-2 among them on lane slots 4, 5
    if (k) {
        int order[5], no = 0;
        for (int t = 0; t < k; t++) if (mf[t] < 3) order[no++] = mf[t];
        for (int t = 0; t < k; t++) if (mf[t] >= 3) order[no++] = mf[t];
        // bits 0-2 beyond the two lane MFMA slots cannot sit on lane slots
        if (lanes3) {
            int low = 0;
            for (int t = 0; t < k; t++) low += mf[t] < 3;
            if (low > 2) return false;
        }
        // keep targets already on MFMA slots (cheaper to read, no other effect)
        int slotOfT[5];
        for (int t = 0; t < k; t++) slotOfT[t] = -1;
        if (keep) {
            for (int t = 0; t < k; t++) {
                const int s = keep->slotOf[order[t]];
                for (int q = 0; q < k; q++)
                    if (s == MF_SLOTS[q] && !slotUsed[s] && !(lanes3 && order[t] < 3 && s >= 6)) {
                        slotOfT[t] = s;
                        slotUsed[s] = true;
                    }
            }
        }
        for (int t = 0; t < k; t++) {
            if (slotOfT[t] >= 0) continue;
            for (int q = 0; q < k; q++)
                if (!slotUsed[MF_SLOTS[q]] && !(lanes3 && order[t] < 3 && MF_SLOTS[q] >= 6)) {
                    slotOfT[t] = MF_SLOTS[q];
                    slotUsed[MF_SLOTS[q]] = true;
                    break;
                }
            if (slotOfT[t] < 0) return false;
        }
        for (int t = 0; t < k; t++) L.place(order[t], slotOfT[t]);
    }
    // bits that must be on lane slots, then bits that must be local; reuse
    // the kept layout's slot where it qualifies
    auto put = [&](uint32_t bits, const int* cands, int nc) {
        for (uint32_t m = bits; m; m &= m - 1) {
            const int b = __builtin_ctz(m);
            if (L.slotOf[b] >= 0) continue;
            int s = keep ? keep->slotOf[b] : -1;
            bool ok = false;
            for (int c = 0; c < nc; c++) ok = ok || cands[c] == s;
            if (!ok || s < 0 || slotUsed[s]) {
                s = -1;
                for (int c = 0; c < nc && s < 0; c++)
                    if (!slotUsed[cands[c]]) s = cands[c];
            }
            if (s < 0) return false;
            slotUsed[s] = true;
            L.place(b, s);
        }
        return true;
    };
    static const int lanes[6] = {0, 1, 2, 3, 4, 5};
    static const int locals[9] = {9, 10, 11, 0, 1, 2, 3, 4, 5};   // register slots first: no exchanges
    if (!put(lane, lanes, 6)) return false;
    if (!put(local, locals, 9)) return false;
    // the rest anywhere: where they were, else the free slots
    static const int any[NS] = {6, 7, 8, 9, 10, 11, 0, 1, 2, 3, 4, 5};
    for (int b = 0; b < NS; b++) {
        if (L.slotOf[b] >= 0) continue;
        int s = keep ? keep->slotOf[b] : -1;
        if (s < 0 || slotUsed[s]) {
            s = -1;
            for (int c = 0; c < NS && s < 0; c++)
                if (!slotUsed[any[c]]) s = any[c];
        }
        slotUsed[s] = true;
        L.place(b, s);
    }
    out = L;
    return true;
}

// the longest window of ops starting at i that one layout serves (at least op i)
size_t window_end(const std::vector<Need>& needs, size_t i, bool lanes3, const Layout* keep, Layout& out) {
    if (!window_layout(needs, i, i, lanes3, keep, out))
        fail(FSA_ERR_INTERNAL, "register-resident pass: an op has no layout");
    size_t j = i;
    Layout cand;
    while (j + 1 < needs.size() && window_layout(needs, i, j + 1, lanes3, keep, cand)) {
        j++;
        out = cand;
    }
    return j;
}

// the remap's permutation: for each new slot, the old slot of the tile bit it holds
TileOp remap_item(const Layout& from, const Layout& to) {
    TileOp t;
    std::memset(&t, 0, sizeof(t));
    t.kind = OP_REMAP;
    t.param = NO_PARAM;
    for (int s = 0; s < NS; s++) t.gmask |= (uint64_t)from.slotOf[to.bitOf[s]] << (4 * s);
    t.gq = to.pack();
    return t;
}

// complex matrix with its bit order permuted: new bit j is old bit src[j]
std::vector<cd> permute_bits(const std::vector<cd>& u, int k, const int* src) {
    const uint64_t D = 1ULL << k;
    auto old_of = [&](uint64_t x) {
        uint64_t o = 0;
        for (int j = 0; j < k; j++)
            if ((x >> j) & 1ULL) o |= 1ULL << src[j];
        return o;
    };
    std::vector<cd> out(D * D);
    for (uint64_t r = 0; r < D; r++)
        for (uint64_t c = 0; c < D; c++) out[r * D + c] = u[old_of(r) * D + old_of(c)];
    return out;
}

}  // namespace

bool plan_rtile(const std::vector<HOp>& ops, const Stage& st, std::vector<TileOp>& items,
                std::vector<cx<double>>& mats, RTilePass& out) {
    if (!st.tiled || st.tq.size() != (size_t)NS) return false;
    int pos[64];
    for (int& p : pos) p = -1;
    uint64_t tileMask = 0;
    for (int i = 0; i < NS; i++) {
        pos[st.tq[i]] = i;
        tileMask |= 1ULL << st.tq[i];
    }
    // tile bits 0-2 must be qubits 0-2 (the contiguous runs of the HBM accesses)
    for (int i = 0; i < 3; i++)
        if (st.tq[i] != (unsigned)i) return false;
    std::vector<Need> needs;
    for (int oi : st.ops) {
        needs.push_back(need_of(ops[oi], pos));
        if (!needs.back().ok) return false;
    }
    const size_t n = needs.size();
    // phase 1: the layout each op runs in: windows of ops that one layout
    // serves, a remap between windows (the first window's layout is the load
    // layout, which needs tile bits 0-2 on lane slots)
    std::vector<Layout> at(n);
    std::vector<char> remap(n, 0);
    Layout load, L;
    // a first op that cannot run with tile bits 0-2 on lane slots (a dense gate
    // on three of them: only two MFMA slots are lane slots) runs after a remap
    // from the identity load layout
    Layout probe;
    bool loadFixed = n == 0 || !window_layout(needs, 0, 0, true, nullptr, probe);
    if (loadFixed) {
        for (int b = 0; b < NS; b++) load.place(b, b);
        L = load;
    }
    size_t i = 0;
    while (i < n) {
        Layout W;
        const bool loadWindow = i == 0 && !loadFixed;
        const size_t j = window_end(needs, i, loadWindow, loadWindow ? nullptr : &L, W);
        if (loadWindow) load = W;
        else remap[i] = 1;
        for (size_t x = i; x <= j; x++) at[x] = W;
        L = W;
        i = j + 1;
    }
    std::vector<Layout> from(n);
    if (n) from[0] = load;
    for (size_t x = 1; x < n; x++) from[x] = at[x - 1];
    // the store needs tile bits 0-2 on lane slots, as the load does
    Layout store = L;
    bool tailRemap = false;
    for (int b = 0; b < 3; b++) {
        if (store.slotOf[b] < 6) continue;
        for (int s = 0; s < 6; s++)
            if (store.bitOf[s] >= 3) { store.swap_slots(store.slotOf[b], s); tailRemap = true; break; }
    }
    // phase 2: items in slot space
    const size_t first = items.size();
    uint32_t matBytes = 0;
    for (size_t i = 0; i < n; i++) {
        if (remap[i]) items.push_back(remap_item(from[i], at[i]));
        const HOp& op = ops[st.ops[i]];
        const Layout& A = at[i];
        TileOp t;
        std::memset(&t, 0, sizeof(t));
        t.kind = op.kind;
        t.param = NO_PARAM;
        switch (op.kind) {
        case OP_MAT1:
            t.p[0] = (uint32_t)A.slotOf[pos[op.q[0]]];
            t.lmask = A.slot_mask(tile_bits_of(op.ctrl, pos));
            t.gmask = op.ctrl & ~tileMask;
            for (int j = 0; j < 4; j++) { t.f[2 * j] = op.mat[j].real(); t.f[2 * j + 1] = op.mat[j].imag(); }
            break;
        case OP_PAULI:
            t.lmask = A.slot_mask(tile_bits_of(op.mxy, pos));
            t.lmask2 = A.slot_mask(tile_bits_of(op.myz, pos));
            t.gmask = op.myz & ~tileMask;
            t.f[0] = op.th.real(); t.f[1] = op.th.imag();
            t.f[2] = op.fe.real(); t.f[3] = op.fe.imag();
            break;
        case OP_PHASE:
            t.lmask = A.slot_mask(tile_bits_of(op.pmask, pos));
            t.gmask = op.pmask & ~tileMask;
            t.f[0] = std::cos(op.theta); t.f[1] = std::sin(op.theta);
            t.f[2] = std::cos(op.theta); t.f[3] = -std::sin(op.theta);
            break;
        case OP_MATK: {
            const int k = (int)op.q.size();
            // matrix bit j of the kernel's order sits on MF_SLOTS[j]: which of the
            // op's targets (its bit i acts on op.q[i]) is that
            int src[5];
            for (int j = 0; j < k; j++) {
                const int b = A.bitOf[MF_SLOTS[j]];
                src[j] = -1;
                for (int q = 0; q < k; q++)
                    if (pos[op.q[q]] == b) src[j] = q;
                if (src[j] < 0) fail(FSA_ERR_INTERNAL, "register-resident pass: MFMA target off its slot");
            }
            const std::vector<cd> u = permute_bits(op.mat, k, src);
            const uint64_t D = 1ULL << k, D2 = 2 * D;
            t.k = (uint32_t)k;
            t.lmask2 = 1;
            t.lmask = matBytes / (uint32_t)sizeof(double);      // LDS offset (reals) of its left half
            matBytes += (uint32_t)(2 * D * D * sizeof(double));
            t.mat_off = (uint32_t)mats.size();
            // real form [[Ur, -Ui], [Ui, Ur]], row-major, two reals per complex slot
            std::vector<double> re(D2 * D2);
            for (uint64_t r = 0; r < D; r++)
                for (uint64_t c = 0; c < D; c++) {
                    const cd x = u[r * D + c];
                    re[r * D2 + c] = x.real();
                    re[r * D2 + D + c] = -x.imag();
                    re[(D + r) * D2 + c] = x.imag();
                    re[(D + r) * D2 + D + c] = x.real();
                }
            for (uint64_t e = 0; e < re.size(); e += 2) mats.push_back(mk<double>(re[e], re[e + 1]));
            break;
        }
        default:
            fail(FSA_ERR_INTERNAL, "register-resident pass: unsupported op");
        }
        items.push_back(t);
    }
    if (tailRemap) items.push_back(remap_item(L, store));
    out.loadMap = load.pack();
    out.storeMap = store.pack();
    out.matBytes = matBytes;
    out.first = first;
    out.count = items.size() - first;
    return true;
}

}  // namespace fsa
