// Host-side symbolic layer: see plan.hpp.
#include "plan.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <numeric>
#include <memory>
#include <set>
#include <unordered_map>
#include <utility>

namespace g2g {

namespace {

// Symmetric adjacency (no self loops, sorted) of a square CSR pattern.
void symmetric_adjacency(int d, const int* Ap, const int* Ai, std::vector<int>& adjp,
                         std::vector<int>& adji)
{
    std::vector<std::vector<int>> nb(static_cast<size_t>(d));
    for (int r = 0; r < d; ++r)
        for (int p = Ap[r]; p < Ap[r + 1]; ++p) {
            const int c = Ai[p];
            if (c < 0 || c >= d) throw std::invalid_argument("CSR column index out of range");
            if (c == r) continue;
            nb[r].push_back(c);
            nb[c].push_back(r);
        }
    adjp.assign(static_cast<size_t>(d) + 1, 0);
    adji.clear();
    for (int r = 0; r < d; ++r) {
        auto& v = nb[r];
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        adji.insert(adji.end(), v.begin(), v.end());
        adjp[r + 1] = static_cast<int>(adji.size());
    }
}

// Column structures of L for a symmetric pattern given as adjacency in pivot
// labels (etree-based symbolic Cholesky).  Returns parent[].
std::vector<int> symbolic_structs(int d, const std::vector<int>& adjp, const std::vector<int>& adji,
                                  std::vector<int>& colp, std::vector<int>& coli)
{
    std::vector<int> parent(static_cast<size_t>(d), -1);
    std::vector<std::vector<int>> children(static_cast<size_t>(d));
    std::vector<int> mark(static_cast<size_t>(d), -1);
    colp.assign(static_cast<size_t>(d) + 1, 0);
    coli.clear();
    std::vector<int> tmp;
    for (int k = 0; k < d; ++k) {
        tmp.clear();
        mark[k] = k;
        for (int p = adjp[k]; p < adjp[k + 1]; ++p) {
            const int i = adji[p];
            if (i > k && mark[i] != k) { mark[i] = k; tmp.push_back(i); }
        }
        for (int c : children[k])
            for (int p = colp[c]; p < colp[c + 1]; ++p) {
                const int i = coli[p];
                if (i > k && mark[i] != k) { mark[i] = k; tmp.push_back(i); }
            }
        std::sort(tmp.begin(), tmp.end());
        coli.insert(coli.end(), tmp.begin(), tmp.end());
        colp[k + 1] = static_cast<int>(coli.size());
        if (!tmp.empty()) {
            parent[k] = tmp.front();
            children[tmp.front()].push_back(k);
        }
    }
    return parent;
}

}  // namespace

int default_walkers(int d)
{
    if (const char* e = std::getenv("G2G_WALKERS")) {
        const int w = std::atoi(e);
        if (w >= 1 && w <= kMaxFactorWalkers) return w;
    }
    return d >= kWalkersMinD ? 2 : 1;
}

int default_bwalkers(int d)
{
    if (const char* e = std::getenv("G2G_BWALKERS")) {
        const int w = std::atoi(e);
        if (w == 1 || w == 2 || w == 4) return w;
    }
    return d >= kWalkersMinD ? 4 : 1;
}

std::vector<int> min_degree_order(int n, const std::vector<int>& adjp, const std::vector<int>& adji,
                                  const std::vector<int>& weight)
{
    std::vector<std::vector<int>> nb(static_cast<size_t>(n));
    for (int v = 0; v < n; ++v) nb[v].assign(adji.begin() + adjp[v], adji.begin() + adjp[v + 1]);
    std::vector<long long> deg(static_cast<size_t>(n), 0);
    auto degree = [&](int v) {
        long long s = 0;
        for (int u : nb[v]) s += weight[u];
        return s;
    };
    std::set<std::pair<long long, int>> pq;
    for (int v = 0; v < n; ++v) {
        deg[v] = degree(v);
        pq.insert({deg[v], v});
    }
    std::vector<char> done(static_cast<size_t>(n), 0);
    std::vector<int> order;
    order.reserve(static_cast<size_t>(n));
    std::vector<int> merged;
    while (!pq.empty()) {
        const int v = pq.begin()->second;
        pq.erase(pq.begin());
        done[v] = 1;
        order.push_back(v);
        const std::vector<int> nv = nb[v];
        for (int u : nv) {
            // N(u) <- N(u) U N(v) \ {u, v}
            merged.clear();
            std::set_union(nb[u].begin(), nb[u].end(), nv.begin(), nv.end(), std::back_inserter(merged));
            nb[u].clear();
            for (int w : merged)
                if (w != u && w != v && !done[w]) nb[u].push_back(w);
            pq.erase({deg[u], u});
            deg[u] = degree(u);
            pq.insert({deg[u], u});
        }
        nb[v].clear();
    }
    return order;
}

std::vector<int> rcm_order(int n, const std::vector<int>& adjp, const std::vector<int>& adji)
{
    std::vector<int> order;
    order.reserve(static_cast<size_t>(n));
    std::vector<char> seen(static_cast<size_t>(n), 0);
    auto deg = [&](int v) { return adjp[v + 1] - adjp[v]; };
    auto bfs_last = [&](int s, std::vector<int>& lvl) {
        lvl.assign(static_cast<size_t>(n), -1);
        std::vector<int> q{s};
        lvl[s] = 0;
        int last = s;
        for (size_t h = 0; h < q.size(); ++h) {
            const int v = q[h];
            last = v;
            for (int p = adjp[v]; p < adjp[v + 1]; ++p)
                if (lvl[adji[p]] < 0) { lvl[adji[p]] = lvl[v] + 1; q.push_back(adji[p]); }
        }
        // among the deepest level, the smallest degree
        int best = last;
        for (int v : q)
            if (lvl[v] == lvl[last] && deg(v) < deg(best)) best = v;
        return best;
    };
    std::vector<int> lvl;
    for (int s0 = 0; s0 < n; ++s0) {
        if (seen[s0]) continue;
        // pseudo-peripheral start of this component
        int s = bfs_last(s0, lvl);
        s = bfs_last(s, lvl);
        std::vector<int> q{s};
        seen[s] = 1;
        std::vector<int> nbr;
        for (size_t h = 0; h < q.size(); ++h) {
            const int v = q[h];
            nbr.clear();
            for (int p = adjp[v]; p < adjp[v + 1]; ++p)
                if (!seen[adji[p]]) { seen[adji[p]] = 1; nbr.push_back(adji[p]); }
            std::sort(nbr.begin(), nbr.end(), [&](int a, int b) { return deg(a) != deg(b) ? deg(a) < deg(b) : a < b; });
            q.insert(q.end(), nbr.begin(), nbr.end());
        }
        order.insert(order.end(), q.begin(), q.end());
    }
    std::reverse(order.begin(), order.end());
    return order;
}

namespace {

// Static slot allocator: lowest free slot first.
struct SlotAlloc {
    std::set<int> free_;
    int hi = 0;
    int get()
    {
        if (!free_.empty()) { int s = *free_.begin(); free_.erase(free_.begin()); return s; }
        return hi++;
    }
    void put(int s) { free_.insert(s); }
};

inline long long key(int i, int j, int d) { return static_cast<long long>(i) * d + j; }

inline int round_up_to(int v, int m) { return (v + m - 1) / m * m; }

// A record of a factorization program: a walk step (single pivot or
// 2-column supernode) or a merge of another walker's pending values.
struct Rec {
    enum { kStep = 0, kMerge = 1, kDump = 2, kGather = 3 };
    int kind = kStep;
    int k = 0, w = 1, c = 0;  // step: first pivot, width, |struct| of its last pivot; dump: first pend index
    std::vector<int> r;       // step: sources then targets (slot | kFirst, -1 none); merge: (dst, src) pairs;
                              // dump: slots; gather: (dst slot, pend index) pairs
};

// One walker's factorization program under construction: its live update
// targets, slot pool, A/U stream positions and records.  Slots a step
// consumes are freed after its targets are placed, so a step's sources and
// targets never share a slot (the kernels read both in one batch).
struct Emitter {
    SymbolicLU& s;
    const std::unordered_map<long long, int>& a_entry;
    bool record = true;  // false: writes no stream table
    // the stream tables a recording emitter fills (the main layout's by default)
    std::vector<int>* apos_tab = &s.a_pos_of_entry;
    std::vector<int>* fpos_tab = &s.f_pos;
    std::vector<int>* uoff_tab = &s.u_off;
    bool count = true;   // adds to nA / n_updates (the main layout only)
    std::unordered_map<long long, int> slot_of;  // live update targets
    std::vector<int> bslot;                      // live forward-substitution targets
    SlotAlloc sa;
    std::vector<int> consumed;
    std::vector<Rec> recs;
    int apos = 0, upos = 0, n_piv = 0;

    Emitter(SymbolicLU& s_, const std::unordered_map<long long, int>& ae)
        : s(s_), a_entry(ae), bslot(static_cast<size_t>(s_.d), -1)
    {
    }
    int take(int i, int j)  // slot of a consumed entry, or -1
    {
        auto it = slot_of.find(key(i, j, s.d));
        if (it == slot_of.end()) return -1;
        const int sl = it->second;
        slot_of.erase(it);
        consumed.push_back(sl);
        return sl;
    }
    void place(int i, int j)  // dense A stream: every entry has a position
    {
        if (record) {
            auto it = a_entry.find(key(i, j, s.d));
            if (it != a_entry.end()) {
                (*apos_tab)[it->second] = apos;
                if (count) ++s.nA;
            }
        }
        ++apos;
    }
    void alloc_target(int i, int j, std::vector<int>& r)
    {
        const long long kk = key(i, j, s.d);
        auto it = slot_of.find(kk);
        if (it == slot_of.end()) {
            const int sl = sa.get();
            slot_of[kk] = sl;
            r.push_back(sl | kFirst);
        } else {
            r.push_back(it->second);
        }
    }
    void alloc_b(int i, std::vector<int>& r)
    {
        if (bslot[i] < 0) {
            bslot[i] = sa.get();
            r.push_back(bslot[i] | kFirst);
        } else {
            r.push_back(bslot[i]);
        }
    }
    void sources(int k, std::vector<int>& r)  // diagonal, b(k), L(struct,k), U(k,struct)
    {
        const int c0 = s.colp[k], c = s.colp[k + 1] - c0;
        if (record) (*fpos_tab)[s.perm[k]] = apos;
        ++apos;
        place(k, k);
        for (int p = c0; p < c0 + c; ++p) place(s.coli[p], k);
        for (int p = c0; p < c0 + c; ++p) place(k, s.coli[p]);
        r.push_back(take(k, k));
        int bref = -1;
        if (bslot[k] >= 0) {
            bref = bslot[k];
            consumed.push_back(bref);
            bslot[k] = -1;
        }
        r.push_back(bref);
        for (int p = c0; p < c0 + c; ++p) r.push_back(take(s.coli[p], k));
        for (int p = c0; p < c0 + c; ++p) r.push_back(take(k, s.coli[p]));
        if (record) (*uoff_tab)[k] = upos;
        upos += c + 2;
    }
    void step(int k)
    {
        const int w = s.step_w[k] == 2 ? 2 : 1;
        const int last = k + w - 1;
        const int l0 = s.colp[last], cl = s.colp[last + 1] - l0;
        Rec st;
        st.k = k;
        st.w = w;
        st.c = cl;
        sources(k, st.r);
        if (w == 2) sources(k + 1, st.r);
        // update targets (i, j) in struct(last) x struct(last), row-major, then b(i)
        for (int pa = l0; pa < l0 + cl; ++pa)
            for (int pb = l0; pb < l0 + cl; ++pb) alloc_target(s.coli[pa], s.coli[pb], st.r);
        for (int p = l0; p < l0 + cl; ++p) alloc_b(s.coli[p], st.r);
        if (record && count) s.n_updates += w == 2 ? 2 * cl * cl : cl * cl;
        for (int sl : consumed) sa.put(sl);
        consumed.clear();
        recs.push_back(std::move(st));
        n_piv += w;
    }
    // appends the records; walker-local slot ids are shifted by `off`, a
    // missing source reads the zero slot
    void encode(std::vector<int>& out, int off, int zero) const
    {
        const auto enc = [](int slot) { return slot * kRefUnit; };
        auto head = [](const Rec& x) {
            return x.kind == Rec::kMerge    ? kMergeHead
                   : x.kind == Rec::kDump   ? kDumpHead
                   : x.kind == Rec::kGather ? kGatherHead
                                            : rec_head(x.c, x.w);
        };
        for (size_t q = 0; q < recs.size(); ++q) {
            const Rec& st = recs[q];
            const int next = q + 1 < recs.size() ? head(recs[q + 1]) : 0;
            const size_t at = out.size();
            if (st.kind == Rec::kDump) {
                const int n = static_cast<int>(st.r.size());
                out.resize(at + static_cast<size_t>(dump_words(n)), enc(zero));
                int* wd = out.data() + at;
                wd[0] = kDumpHead | (next << 16);
                wd[1] = n;
                wd[2] = st.k;
                for (int i = 0; i < n; ++i) wd[3 + i] = enc(st.r[i] + off);
                continue;
            }
            if (st.kind == Rec::kGather) {
                const int n = static_cast<int>(st.r.size() / 2);
                out.resize(at + static_cast<size_t>(pair_rec_words(n)), enc(zero));
                int* wd = out.data() + at;
                wd[0] = kGatherHead | (next << 16);
                wd[1] = n;
                for (int i = 0; i < n; ++i) {
                    wd[2 + 2 * i] = enc(st.r[2 * i] + off);
                    wd[3 + 2 * i] = st.r[2 * i + 1];
                }
                continue;
            }
            if (st.kind == Rec::kMerge) {
                const int n = static_cast<int>(st.r.size() / 2);
                out.resize(at + static_cast<size_t>(mrec_words(n)), enc(zero));
                int* wd = out.data() + at;
                wd[0] = kMergeHead | (next << 16);
                wd[1] = n;
                for (int i = 0; i < 2 * n; ++i) wd[2 + i] = enc(st.r[i]);
                continue;
            }
            const int c = st.c;
            const int ns = st.w == 2 ? frec2_src_words(c) : frec_src_words(c);
            out.resize(at + static_cast<size_t>(st.w == 2 ? frec2_words(c) : frec_words(c)), enc(zero));
            int* wd = out.data() + at;
            wd[0] = rec_head(c, st.w) | (next << 16);
            const int nsrc = ns - 1;
            for (int q2 = 0; q2 < nsrc; ++q2) wd[1 + q2] = enc(st.r[q2] < 0 ? zero : (st.r[q2] & kSlotMask) + off);
            for (int i = 0; i < c * c + c; ++i) wd[ns + i] = enc((st.r[nsrc + i] & kSlotMask) + off);
        }
    }
};

std::vector<int> partition_steps(const SymbolicLU& s, const std::vector<int>& parent,
                                 const std::vector<int>& step_first, int W);
void build_backward(SymbolicLU& s, const std::vector<int>& group, int W);

// where each step of the main programs landed, and each split set's range
struct SplitLayout {
    int n = 0;                      // programs: G * F fine sets, G middles, the top (0: no split walk)
    int n_fine = 0, n_mid = 0;      // G * F fine sets, then G middle parts, then the top part
    std::vector<int> fine;          // per pivot: its program
};
void build_split(SymbolicLU& s, const SplitLayout& L, const std::vector<int>& step_first,
                 const std::unordered_map<long long, int>& a_entry);

}  // namespace

SymbolicLU build_symbolic(int d, const int* Ap, const int* Ai, std::vector<int> order, int walkers, int bwalkers)
{
    if (d <= 0) throw std::invalid_argument("matrix dimension must be positive");
    SymbolicLU s;
    s.d = d;
    std::vector<int> adjp, adji;
    symmetric_adjacency(d, Ap, Ai, adjp, adji);
    if (order.empty()) order = min_degree_order(d, adjp, adji, std::vector<int>(static_cast<size_t>(d), 1));
    if (static_cast<int>(order.size()) != d) throw std::invalid_argument("order has the wrong length");
    {
        std::vector<char> seen(static_cast<size_t>(d), 0);
        for (int v : order) {
            if (v < 0 || v >= d || seen[v]) throw std::invalid_argument("order is not a permutation");
            seen[v] = 1;
        }
    }
    // adjacency in the labels of `order`
    auto relabel = [&](const std::vector<int>& perm, std::vector<int>& pp, std::vector<int>& pi) {
        std::vector<int> inv(static_cast<size_t>(d));
        for (int k = 0; k < d; ++k) inv[perm[k]] = k;
        pp.assign(static_cast<size_t>(d) + 1, 0);
        pi.clear();
        std::vector<int> row;
        for (int k = 0; k < d; ++k) {
            const int v = perm[k];
            row.clear();
            for (int p = adjp[v]; p < adjp[v + 1]; ++p) row.push_back(inv[adji[p]]);
            std::sort(row.begin(), row.end());
            pi.insert(pi.end(), row.begin(), row.end());
            pp[k + 1] = static_cast<int>(pi.size());
        }
    };
    std::vector<int> pp, pi, colp, coli;
    relabel(order, pp, pi);
    std::vector<int> parent = symbolic_structs(d, pp, pi, colp, coli);
    // postorder of the elimination forest.  Children are visited in
    // decreasing (peak - out) order (Liu's rule for the smallest working
    // stack): out_c = c_c^2 + c_c live update/rhs slots a finished subtree
    // leaves behind, peak_c the most it holds while running.
    std::vector<std::vector<int>> children(static_cast<size_t>(d));
    std::vector<int> roots;
    for (int k = 0; k < d; ++k) (parent[k] >= 0 ? children[parent[k]] : roots).push_back(k);
    {
        std::vector<long long> out(static_cast<size_t>(d)), peak(static_cast<size_t>(d));
        for (int k = 0; k < d; ++k) {  // children have smaller labels than parents
            const long long ck = colp[k + 1] - colp[k];
            out[k] = ck * ck + ck;
            auto& ch = children[k];
            std::sort(ch.begin(), ch.end(), [&](int a, int b) {
                const long long x = peak[a] - out[a], y = peak[b] - out[b];
                return x != y ? x > y : a < b;
            });
            long long held = 0, pk = 0;
            for (int c : ch) {
                pk = std::max(pk, held + peak[c]);
                held += out[c];
            }
            peak[k] = std::max(pk, held + out[k]);
        }
        std::sort(roots.begin(), roots.end(), [&](int a, int b) {
            const long long x = peak[a] - out[a], y = peak[b] - out[b];
            return x != y ? x > y : a < b;
        });
    }
    std::vector<int> post;
    post.reserve(static_cast<size_t>(d));
    for (int r : roots) {
        std::vector<std::pair<int, size_t>> st{{r, 0}};
        while (!st.empty()) {
            auto& top = st.back();
            if (top.second < children[top.first].size()) {
                const int c = children[top.first][top.second++];
                st.push_back({c, 0});
            } else {
                post.push_back(top.first);
                st.pop_back();
            }
        }
    }
    s.perm.resize(static_cast<size_t>(d));
    for (int k = 0; k < d; ++k) s.perm[k] = order[post[k]];
    s.iperm.resize(static_cast<size_t>(d));
    for (int k = 0; k < d; ++k) s.iperm[s.perm[k]] = k;
    relabel(s.perm, pp, pi);
    parent = symbolic_structs(d, pp, pi, s.colp, s.coli);
    {
        std::vector<int> h(static_cast<size_t>(d), 0);
        for (int k = 0; k < d; ++k)
            if (parent[k] >= 0) h[parent[k]] = std::max(h[parent[k]], h[k] + 1);
        s.height = d ? 1 + *std::max_element(h.begin(), h.end()) : 0;
    }
    s.nLU = d + 2 * static_cast<int>(s.coli.size());

    // which (row, col) pivot-label entries are present in the input A
    std::unordered_map<long long, int> a_entry;  // key -> input CSR position
    a_entry.reserve(static_cast<size_t>(Ap[d]) * 2);
    for (int r = 0; r < d; ++r)
        for (int p = Ap[r]; p < Ap[r + 1]; ++p)
            a_entry[key(s.iperm[r], s.iperm[Ai[p]], d)] = p;
    s.a_pos_of_entry.assign(static_cast<size_t>(Ap[d]), -1);
    s.f_pos.assign(static_cast<size_t>(d), -1);
    s.nA = 0;

    // Steps of the walk: a single pivot, or a 2-column supernode (k, k+1) when
    // struct(k) = {k+1} + struct(k+1) -- the P and Q unknowns of one bus,
    // eliminated together.  A supernode reads the sources of both pivots,
    // applies pivot k to pivot k+1's row and column in registers, and updates
    // struct(k+1) x struct(k+1) once with a rank-2 term.  The A and U streams
    // keep their per-pivot layout.
    s.step_w.assign(static_cast<size_t>(d), 1);
    std::vector<int> step_first;
    s.max_c = 0;
    for (int k = 0; k < d;) {
        const int c0 = s.colp[k], c = s.colp[k + 1] - c0;
        s.max_c = std::max(s.max_c, c);
        bool pair = false;
        if (k + 1 < d && c >= 1 && s.coli[c0] == k + 1) {
            const int c1 = s.colp[k + 2] - s.colp[k + 1];
            pair = c1 == c - 1 && c1 <= kFastC2 &&
                   std::equal(s.coli.begin() + c0 + 1, s.coli.begin() + c0 + c, s.coli.begin() + s.colp[k + 1]);
        }
        step_first.push_back(k);
        if (pair) {
            s.step_w[k] = 2;
            s.step_w[k + 1] = 0;
            s.max_c = std::max(s.max_c, c - 1);
            k += 2;
        } else {
            k += 1;
        }
    }
    s.n_steps = static_cast<int>(step_first.size());
    s.u_off.assign(static_cast<size_t>(d) + 1, 0);
    const int W = std::max(1, std::min(walkers, kMaxFactorWalkers));
    // G subtree sets (the backward walkers), G / W per factorization walker;
    // group G is the top part
    int G = bwalkers > 0 ? std::min(bwalkers, kMaxWalkers) : W;
    if (G < W || G % W != 0) G = W;
    // The walker sets: G sets of whole subtrees (largest first into the
    // lightest) below a small top part.
    const std::vector<int> group = partition_steps(s, parent, step_first, G);
    // Split walk (large matrices): each walker set is cut once more into F
    // fine sets of whole subtrees and a middle part above them; the split
    // walk has stream positions of its own (build_split).
    SplitLayout SL;
    if (d >= kSplitMinD) {
        const int F = std::max(2, kSplitSets / G);
        SL.n_fine = G * F;
        SL.n_mid = G;
        SL.n = SL.n_fine + SL.n_mid + 1;
        SL.fine.assign(static_cast<size_t>(d), SL.n - 1);
        for (int g = 0; g < G; ++g) {
            std::vector<int> sub;
            for (int k : step_first)
                if (group[k] == g) sub.push_back(k);
            if (sub.empty()) continue;
            const std::vector<int> fg = partition_steps(s, parent, sub, F);  // F: the middle part
            for (int k : sub) {
                const int f = fg[k];
                const int prog = f == F ? SL.n_fine + g : g * F + f;
                SL.fine[k] = prog;
                if (s.step_w[k] == 2) SL.fine[k + 1] = prog;
            }
        }
        for (int k = 0; k < d; ++k)
            if (group[k] == G) SL.fine[k] = SL.n - 1;
    }
    const int gpw = G / W;
    // U-stream range and pivots of every set (and of the top part)
    std::vector<int> gu_off(static_cast<size_t>(G) + 1, 0), gu_end(static_cast<size_t>(G) + 1, 0),
        gpiv(static_cast<size_t>(G) + 1, 0);

    std::vector<std::unique_ptr<Emitter>> em;
    for (int w = 0; w < W; ++w) em.push_back(std::make_unique<Emitter>(s, a_entry));
    Emitter& e0 = *em[0];
    // first phase: walkers 1 .. W-1 lay out their segments first, each at an
    // aligned offset, then walker 0's (whose segment continues with the top part)
    std::vector<int> a_off(static_cast<size_t>(W), 0), u_off(static_cast<size_t>(W), 0);
    int a_cur = 0, u_cur = 0;
    // a walker walks its sets last to first, so its set of lowest index lies
    // right before whatever follows in its U segment (walker 0: the top part)
    auto walk_sets = [&](Emitter& e, int w) {
        for (int g = (w + 1) * gpw - 1; g >= w * gpw; --g) {
            gu_off[g] = e.upos;
            const int p0 = e.n_piv;
            for (int k : step_first)
                if (group[k] == g) e.step(k);
            gu_end[g] = e.upos;
            gpiv[g] = e.n_piv - p0;
        }
    };
    for (int w = 1; w < W; ++w) {
        a_off[w] = a_cur;
        u_off[w] = u_cur;
        em[w]->apos = a_cur;
        em[w]->upos = u_cur;
        walk_sets(*em[w], w);
        a_cur = round_up_to(em[w]->apos, kSegAlign);
        u_cur = round_up_to(em[w]->upos, kSegAlign);
    }
    a_off[0] = a_cur;
    u_off[0] = u_cur;
    e0.apos = a_cur;
    e0.upos = u_cur;
    walk_sets(e0, 0);
    const int n_phase1_0 = e0.n_piv;
    // walker w's slots sit after walker 0's and those of walkers 1 .. w-1
    std::vector<int> slot_base(static_cast<size_t>(W), 0);
    if (W > 1) {
        slot_base[1] = e0.sa.hi;
        for (int w = 2; w < W; ++w) slot_base[w] = slot_base[w - 1] + em[w - 1]->sa.hi;
        // merge the other walkers' pending contributions into walker 0's
        // numbering: an entry already held there is summed by a merge record,
        // an entry held nowhere yet keeps the walker's slot
        std::vector<int> merges;
        std::set<int> adopted;
        for (int w = 1; w < W; ++w) {
            Emitter& ew = *em[w];
            for (const auto& kv : ew.slot_of) {
                const int g = kv.second + slot_base[w];
                auto it = e0.slot_of.find(kv.first);
                if (it != e0.slot_of.end()) {
                    merges.push_back(it->second);
                    merges.push_back(g);
                } else {
                    e0.slot_of[kv.first] = g;
                    adopted.insert(g);
                }
            }
            for (int i = 0; i < d; ++i) {
                if (ew.bslot[i] < 0) continue;
                const int g = ew.bslot[i] + slot_base[w];
                if (e0.bslot[i] >= 0) {
                    merges.push_back(e0.bslot[i]);
                    merges.push_back(g);
                } else {
                    e0.bslot[i] = g;
                    adopted.insert(g);
                }
            }
            ew.slot_of.clear();
        }
        // the merge records (at least one: it carries the wait for the others)
        const size_t npairs = merges.size() / 2;
        for (size_t q = 0; q == 0 || q < npairs; q += kMergePairs) {
            Rec m;
            m.kind = Rec::kMerge;
            const size_t q1 = std::min(npairs, q + kMergePairs);
            m.r.assign(merges.begin() + 2 * q, merges.begin() + 2 * q1);
            e0.recs.push_back(std::move(m));
            if (npairs == 0) break;
        }
        // the top part allocates from every slot not holding a live value
        const int hi = slot_base[W - 1] + em[W - 1]->sa.hi;
        for (int j = slot_base[1]; j < hi; ++j)
            if (!adopted.count(j)) e0.sa.put(j);
        e0.sa.hi = hi;
    }
    // the top part (walker 0, after the merge)
    gu_off[G] = e0.upos;
    for (int k : step_first)
        if (group[k] == G) e0.step(k);
    gu_end[G] = e0.upos;
    gpiv[G] = e0.n_piv - n_phase1_0;
    int covered = 0;
    for (int w = 0; w < W; ++w) covered += em[w]->n_piv;
    if (covered != d) throw std::logic_error("symbolic: walkers do not cover every pivot");
    if (!e0.slot_of.empty()) throw std::logic_error("symbolic: update targets left unconsumed");
    for (int i = 0; i < d; ++i)
        if (e0.bslot[i] >= 0) throw std::logic_error("symbolic: forward-substitution targets left unconsumed");
    s.slots_f = e0.sa.hi;
    s.a_len = e0.apos;
    s.u_len = e0.upos;
    s.u_off[d] = s.u_len;
    s.walkers = W;
    s.n_top = gpiv[G];
    s.bwalkers = G;
    s.bseg.assign(static_cast<size_t>(G), SymbolicLU::Seg{});
    for (int g = 0; g < G; ++g) {
        s.bseg[g].u_off = gu_off[g];
        s.bseg[g].u_len = gu_end[g] - gu_off[g];
        s.bseg[g].n_piv = gpiv[g];
    }
    // set 0's walker also walks the top part, which follows set 0 in the U stream
    if (gu_end[0] != gu_off[G]) throw std::logic_error("symbolic: the top part does not follow set 0");
    s.bseg[0].u_len = gu_end[G] - gu_off[0];
    s.bseg[0].n_piv += gpiv[G];
    s.seg.assign(static_cast<size_t>(W), SymbolicLU::Seg{});
    for (int w = 0; w < W; ++w) {
        s.seg[w].a_off = a_off[w];
        s.seg[w].a_len = em[w]->apos - a_off[w];
        s.seg[w].u_off = u_off[w];
        s.seg[w].u_len = em[w]->upos - u_off[w];
        s.seg[w].n_piv = em[w]->n_piv;
    }
    // encode: refs as byte offsets; a missing source reads the zero slot
    s.prog_f.clear();
    for (int w = 0; w < W; ++w) {
        s.seg[w].prog_f_off = static_cast<int>(s.prog_f.size());
        em[w]->encode(s.prog_f, slot_base[w], s.slots_f);
        s.prog_f.resize(static_cast<size_t>(round_up_to(static_cast<int>(s.prog_f.size()), kProgChunk)),
                        s.slots_f * kRefUnit);
        s.seg[w].prog_f_len = static_cast<int>(s.prog_f.size()) - s.seg[w].prog_f_off;
    }
    build_backward(s, group, G);
    if (SL.n) build_split(s, SL, step_first, a_entry);
    return s;
}

namespace {

// Cuts the elimination forest (at step granularity) into W groups of whole
// subtrees plus a top part (group W) and returns the group of every pivot.
// A threshold T puts every step whose subtree costs more than T into the top;
// the remaining maximal subtrees are packed into W bins, largest first into
// the lightest bin.  The T of the smallest makespan (largest bin + top) wins.
std::vector<int> partition_steps(const SymbolicLU& s, const std::vector<int>& parent,
                                 const std::vector<int>& step_first, int W)
{
    const int d = s.d;
    std::vector<int> group(static_cast<size_t>(d), 0);
    if (W <= 1) return group;
    const int ns = static_cast<int>(step_first.size());
    std::vector<int> step_of(static_cast<size_t>(d), -1);  // -1: a step outside the given ones
    for (int q = 0; q < ns; ++q) {
        const int k = step_first[q];
        step_of[k] = q;
        if (s.step_w[k] == 2) step_of[k + 1] = q;
    }
    std::vector<int> sparent(static_cast<size_t>(ns), -1);
    std::vector<double> cost(static_cast<size_t>(ns)), sub(static_cast<size_t>(ns));
    for (int q = 0; q < ns; ++q) {
        const int k = step_first[q];
        const int last = s.step_w[k] == 2 ? k + 1 : k;
        const int c = s.colp[last + 1] - s.colp[last];
        cost[q] = 3.0 + c + 0.25 * c * c;  // a step's walk time, roughly
        if (parent[last] >= 0) sparent[q] = step_of[parent[last]];  // -1 when the parent is not given
    }
    for (int q = 0; q < ns; ++q) sub[q] += cost[q];  // children precede parents
    for (int q = 0; q < ns; ++q)
        if (sparent[q] >= 0) sub[sparent[q]] += sub[q];
    std::vector<double> cand(sub);
    std::sort(cand.begin(), cand.end());
    cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
    double best = 1e300, bestT = cand.back();
    const size_t stride = std::max<size_t>(1, cand.size() / 512);
    for (size_t ci = 0; ci < cand.size(); ci += stride) {
        const double T = cand[ci];
        double top = 0.0;
        std::vector<double> parts;
        for (int q = 0; q < ns; ++q) {
            if (sub[q] > T) {
                top += cost[q];
            } else if (sparent[q] < 0 || sub[sparent[q]] > T) {
                parts.push_back(sub[q]);
            }
        }
        std::sort(parts.begin(), parts.end(), std::greater<double>());
        std::vector<double> bins(static_cast<size_t>(W), 0.0);
        for (double x : parts) *std::min_element(bins.begin(), bins.end()) += x;
        const double mk = *std::max_element(bins.begin(), bins.end()) + top;
        if (mk < best) { best = mk; bestT = T; }
    }
    // assign: part roots largest first into the lightest bin, then propagate
    std::vector<int> sg(static_cast<size_t>(ns), -1);
    std::vector<std::pair<double, int>> roots;
    for (int q = 0; q < ns; ++q) {
        if (sub[q] > bestT) sg[q] = W;
        else if (sparent[q] < 0 || sub[sparent[q]] > bestT) roots.push_back({sub[q], q});
    }
    std::sort(roots.begin(), roots.end(), [](const auto& a, const auto& b) {
        return a.first != b.first ? a.first > b.first : a.second < b.second;
    });
    std::vector<double> bins(static_cast<size_t>(W), 0.0);
    for (const auto& rt : roots) {
        const int b = static_cast<int>(std::min_element(bins.begin(), bins.end()) - bins.begin());
        bins[b] += rt.first;
        sg[rt.second] = b;
    }
    for (int q = ns - 1; q >= 0; --q)  // parents precede children in reverse order
        if (sg[q] < 0) sg[q] = sg[sparent[q]];
    for (int k = 0; k < d; ++k)
        if (step_of[k] >= 0) group[k] = sg[step_of[k]];
    return group;
}

// Backward programs, one per walker, records in processing order: walker 0
// walks the top part (last pivot first) and then its subtrees; walker w > 0
// waits for the top and walks its subtrees.  Header word (c | next << 16),
// perm[k], then the refs of x(j), j in struct_k, and the slot x(k) is kept in
// (the trash slot when no later step reads it).  The x of a top pivot is read
// by every walker and keeps its slot to the end.
void build_backward(SymbolicLU& s, const std::vector<int>& group, int W)
{
    const int d = s.d;
    std::vector<int> last(static_cast<size_t>(d), d);  // smallest pivot reading x(j)
    for (int k = 0; k < d; ++k)
        for (int p = s.colp[k]; p < s.colp[k + 1]; ++p) last[s.coli[p]] = std::min(last[s.coli[p]], k);
    struct BRec {
        int k, w, c;  // first pivot of the step, width, |struct| of its last pivot
        std::vector<int> r;
    };
    std::vector<std::vector<BRec>> recs(static_cast<size_t>(W));
    std::vector<int> xslot(static_cast<size_t>(d), -1);
    // pools: the top part's x slots first, then one pool per walker
    SlotAlloc top_sa;
    std::vector<SlotAlloc> pool(static_cast<size_t>(W));
    auto walk = [&](int wk, bool top_phase, SlotAlloc& sa, int base) {
        for (int kk = d - 1; kk >= 0; --kk) {
            const bool pair = s.step_w[kk] == 0;  // kk is the second pivot of a pair
            const int k = pair ? kk - 1 : kk;
            kk = k;
            const int g = group[k];
            if (top_phase ? g != W : g != wk) continue;
            const int last_p = pair ? k + 1 : k;
            const int c0 = s.colp[last_p], c = s.colp[last_p + 1] - c0;
            BRec rec{k, pair ? 2 : 1, c, {}};
            std::vector<int> to_free;
            for (int p = c0; p < c0 + c; ++p) {
                const int j = s.coli[p];
                if (xslot[j] < 0) throw std::logic_error("backward: x(j) read before it is computed");
                rec.r.push_back(xslot[j]);
                if (last[j] == k && group[j] != W) to_free.push_back(xslot[j]);
            }
            // every read of this step happens before its results are stored,
            // so a slot freed here may hold one of them
            for (int sl : to_free) sa.put(sl - base);
            auto out_slot = [&](int j) {
                int out = -1;
                if (last[j] < k) {
                    out = sa.get() + base;
                    xslot[j] = out;
                }
                rec.r.push_back(out);
            };
            if (pair) out_slot(k + 1);  // x(k+1): read by k in registers; a slot only for earlier pivots
            out_slot(k);
            recs[static_cast<size_t>(wk)].push_back(std::move(rec));
        }
    };
    walk(0, true, top_sa, 0);
    int base = top_sa.hi;
    for (int w = 0; w < W; ++w) {
        walk(w, false, pool[w], base);
        base += pool[w].hi;
    }
    s.slots_b = base;
    const auto enc = [&](int slot) { return (slot < 0 ? s.slots_b + 1 : slot) * kRefUnit; };
    s.prog_b.clear();
    for (int w = 0; w < W; ++w) {
        s.bseg[w].prog_b_off = static_cast<int>(s.prog_b.size());
        const auto& rv = recs[static_cast<size_t>(w)];
        for (size_t q = 0; q < rv.size(); ++q) {
            const BRec& rec = rv[q];
            const int c = rec.c;
            const int next = q + 1 < rv.size() ? rec_head(rv[q + 1].c, rv[q + 1].w) : 0;
            const size_t at = s.prog_b.size();
            const int nw = rec.w == 2 ? brec2_words(c) : brec_words(c);
            s.prog_b.resize(at + static_cast<size_t>(nw), enc(-1));
            int* wd = s.prog_b.data() + at;
            wd[0] = rec_head(c, rec.w) | (next << 16);
            if (rec.w == 2) {
                wd[1] = s.perm[rec.k + 1];
                wd[2] = s.perm[rec.k];
                for (size_t i = 0; i < rec.r.size(); ++i) wd[3 + i] = enc(rec.r[i]);
            } else {
                wd[1] = s.perm[rec.k];
                for (size_t i = 0; i < rec.r.size(); ++i) wd[2 + i] = enc(rec.r[i]);
            }
        }
        s.prog_b.resize(static_cast<size_t>(round_up_to(static_cast<int>(s.prog_b.size()), kProgChunk)), enc(-1));
        s.bseg[w].prog_b_len = static_cast<int>(s.prog_b.size()) - s.bseg[w].prog_b_off;
    }
}

// The split walk's programs (see SymbolicLU): fine sets (level 0), middle
// parts (level 1), the top part (level 2), each with its own slot numbering
// (all sharing the zero and trash slots slots_fs / slots_fs + 1), laid out
// at the A / U positions the main programs gave the same steps.  A program
// dumps the update targets it leaves pending; a higher program gathers each
// pending value right before the step that consumes it, so its slots stay
// as few as the main walk's.
int split_level(const SplitLayout& L, int j) { return j < L.n_fine ? 0 : j < L.n_fine + L.n_mid ? 1 : 2; }

void build_split(SymbolicLU& s, const SplitLayout& L, const std::vector<int>& step_first,
                 const std::unordered_map<long long, int>& a_entry)
{
    const int d = s.d, n = L.n;
    s.n_split = n;
    s.split_fine = L.n_fine;
    s.split_mid = L.n_mid;
    // the split layout: the programs' A / U ranges one after the other, each
    // starting at a multiple of kSegAlign positions
    s.split_apos.assign(s.a_pos_of_entry.size(), -1);
    s.split_fpos.assign(static_cast<size_t>(d), -1);
    std::vector<int> split_uoff(static_cast<size_t>(d) + 1, 0);
    std::vector<int> a0(static_cast<size_t>(n)), a1(static_cast<size_t>(n)), u0(static_cast<size_t>(n)),
        u1(static_cast<size_t>(n)), piv(static_cast<size_t>(n));
    int a_cur = 0, u_cur = 0;
    // ---- factorization ----
    std::vector<std::unique_ptr<Emitter>> em;
    const long long bkey0 = static_cast<long long>(d) * d;  // keys of forward-substitution targets: bkey0 + i
    std::unordered_map<long long, std::vector<int>> pending;  // key -> pend indices not gathered yet
    int pend_len = 0;
    for (int j = 0; j < n; ++j) {
        em.push_back(std::make_unique<Emitter>(s, a_entry));
        Emitter& e = *em.back();
        e.apos_tab = &s.split_apos;
        e.fpos_tab = &s.split_fpos;
        e.uoff_tab = &split_uoff;
        e.count = false;
        e.apos = a_cur;
        e.upos = u_cur;
        a0[j] = a_cur;
        u0[j] = u_cur;
        for (int k : step_first) {
            if (L.fine[k] != j) continue;
            // the pending values this step consumes as sources, gathered first
            std::vector<int> gat;
            auto gather = [&](long long kk, bool bk, int i) {
                auto it = pending.find(kk);
                if (it == pending.end()) return;
                int dst;
                if (bk) {
                    if (e.bslot[i] < 0) e.bslot[i] = e.sa.get();
                    dst = e.bslot[i];
                } else {
                    auto jt = e.slot_of.find(kk);
                    if (jt == e.slot_of.end()) jt = e.slot_of.emplace(kk, e.sa.get()).first;
                    dst = jt->second;
                }
                for (int q : it->second) {
                    gat.push_back(dst);
                    gat.push_back(q);
                }
                pending.erase(it);
            };
            const int w = s.step_w[k] == 2 ? 2 : 1;
            for (int pv = k; pv < k + w; ++pv) {
                gather(key(pv, pv, d), false, 0);
                gather(bkey0 + pv, true, pv);
                for (int p = s.colp[pv]; p < s.colp[pv + 1]; ++p) {
                    gather(key(s.coli[p], pv, d), false, 0);
                    gather(key(pv, s.coli[p], d), false, 0);
                }
            }
            for (size_t q = 0; q < gat.size(); q += 2 * kSplitRecItems) {
                Rec r;
                r.kind = Rec::kGather;
                r.r.assign(gat.begin() + static_cast<long>(q),
                           gat.begin() + static_cast<long>(std::min(gat.size(), q + 2 * kSplitRecItems)));
                e.recs.push_back(std::move(r));
            }
            e.step(k);
        }
        a1[j] = e.apos;
        u1[j] = e.upos;
        piv[j] = e.n_piv;
        a_cur = round_up_to(e.apos, kSegAlign);
        u_cur = round_up_to(e.upos, kSegAlign);
        if (j == n - 1) break;
        // what this program leaves pending goes to the pend buffer
        std::vector<std::pair<long long, int>> pv;
        for (const auto& kv : e.slot_of) pv.push_back(kv);
        for (int i = 0; i < d; ++i)
            if (e.bslot[i] >= 0) pv.push_back({bkey0 + i, e.bslot[i]});
        std::sort(pv.begin(), pv.end());
        for (size_t q = 0; q < pv.size(); q += kSplitRecItems) {
            Rec r;
            r.kind = Rec::kDump;
            r.k = pend_len + static_cast<int>(q);
            for (size_t i = q; i < std::min(pv.size(), q + kSplitRecItems); ++i) r.r.push_back(pv[i].second);
            e.recs.push_back(std::move(r));
        }
        for (size_t q = 0; q < pv.size(); ++q) pending[pv[q].first].push_back(pend_len + static_cast<int>(q));
        pend_len += static_cast<int>(pv.size());
        e.slot_of.clear();
    }
    s.pend_len = pend_len;
    s.split_a_len = a1[n - 1];
    s.split_u_len = u1[n - 1];
    int covered = 0;
    for (int j = 0; j < n; ++j) covered += piv[j];
    if (covered != d) throw std::logic_error("split: the programs do not cover every pivot");
    for (int q : s.split_fpos)
        if (q < 0) throw std::logic_error("split: an equation without a right-hand side position");
    Emitter& et = *em.back();
    if (!pending.empty()) throw std::logic_error("split: pending values never gathered");
    if (!et.slot_of.empty()) throw std::logic_error("split: update targets left unconsumed");
    for (int i = 0; i < d; ++i)
        if (et.bslot[i] >= 0) throw std::logic_error("split: forward-substitution targets left unconsumed");
    int S = 0;
    for (const auto& e : em) S = std::max(S, e->sa.hi);
    s.slots_fs = S;
    s.prog_fs.clear();
    s.split_seg.assign(static_cast<size_t>(n), SymbolicLU::Seg{});
    for (int j = 0; j < n; ++j) {
        SymbolicLU::Seg& g = s.split_seg[j];
        g.prog_f_off = static_cast<int>(s.prog_fs.size());
        em[j]->encode(s.prog_fs, 0, S);
        s.prog_fs.resize(static_cast<size_t>(round_up_to(static_cast<int>(s.prog_fs.size()), kProgChunk)), S * kRefUnit);
        g.prog_f_len = static_cast<int>(s.prog_fs.size()) - g.prog_f_off;
        g.a_off = a0[j];
        g.a_len = a1[j] - a0[j];
        g.u_off = u0[j];
        g.u_len = u1[j] - u0[j];
        g.n_piv = piv[j];
    }

    // ---- backward substitution: the top part, then the middle parts, then the fine sets ----
    std::vector<int> set_of(static_cast<size_t>(d));
    for (int k = 0; k < d; ++k) set_of[k] = L.fine[s.step_w[k] == 0 ? k - 1 : k];
    struct BRec {
        int kind, k, w, c;  // kind 0: step, 1: preload (r = (dst, x index) pairs)
        std::vector<int> r;
    };
    std::vector<std::vector<BRec>> brecs(static_cast<size_t>(n));
    std::vector<int> bslots(static_cast<size_t>(n), 0);
    for (int j = 0; j < n; ++j) {
        // the last pivot of this program reading x(i)
        std::vector<int> last(static_cast<size_t>(d), d);
        std::vector<char> need(static_cast<size_t>(d), 0);  // top x a set reads
        for (int k = 0; k < d; ++k) {
            if (set_of[k] != j) continue;
            for (int p = s.colp[k]; p < s.colp[k + 1]; ++p) {
                const int i = s.coli[p];
                last[i] = std::min(last[i], k);
                if (set_of[i] != j) {  // an ancestor's x, solved by a program of a higher level
                    if (split_level(L, set_of[i]) <= split_level(L, j))
                        throw std::logic_error("split: a program reads the x of a program of its own level");
                    need[i] = 1;
                }
            }
        }
        SlotAlloc sa;
        std::vector<int> xslot(static_cast<size_t>(d), -1);
        auto& rv = brecs[static_cast<size_t>(j)];
        std::vector<int> pre;
        for (int i = 0; i < d; ++i)
            if (need[i]) {
                xslot[i] = sa.get();
                pre.push_back(xslot[i]);
                pre.push_back(s.perm[i]);
            }
        for (size_t q = 0; q < pre.size(); q += 2 * kSplitRecItems) {
            BRec r{1, 0, 0, 0, {}};
            r.r.assign(pre.begin() + static_cast<long>(q), pre.begin() + static_cast<long>(std::min(pre.size(), q + 2 * kSplitRecItems)));
            rv.push_back(std::move(r));
        }
        int npiv = 0;
        for (int kk = d - 1; kk >= 0; --kk) {
            const bool pair = s.step_w[kk] == 0;
            const int k = pair ? kk - 1 : kk;
            kk = k;
            if (set_of[k] != j) continue;
            const int last_p = pair ? k + 1 : k;
            const int c0 = s.colp[last_p], c = s.colp[last_p + 1] - c0;
            BRec rec{0, k, pair ? 2 : 1, c, {}};
            std::vector<int> to_free;
            for (int p = c0; p < c0 + c; ++p) {
                const int i = s.coli[p];
                if (xslot[i] < 0) throw std::logic_error("split backward: x(i) read before it is computed");
                rec.r.push_back(xslot[i]);
                if (last[i] == k) to_free.push_back(xslot[i]);
            }
            for (int sl : to_free) sa.put(sl);
            auto out_slot = [&](int i) {
                int out = -1;
                if (last[i] < k) {
                    out = sa.get();
                    xslot[i] = out;
                }
                rec.r.push_back(out);
            };
            if (pair) out_slot(k + 1);
            out_slot(k);
            rv.push_back(std::move(rec));
            npiv += pair ? 2 : 1;
        }
        if (npiv != piv[j]) throw std::logic_error("split backward: pivot count");
        bslots[j] = sa.hi;
    }
    const int SB = *std::max_element(bslots.begin(), bslots.end());
    s.slots_bs = SB;
    const auto enc = [&](int slot) { return (slot < 0 ? SB + 1 : slot) * kRefUnit; };
    auto head = [](const BRec& r) { return r.kind == 1 ? kPreloadHead : rec_head(r.c, r.w); };
    s.prog_bs.clear();
    s.split_bseg.assign(static_cast<size_t>(n), SymbolicLU::Seg{});
    for (int j = 0; j < n; ++j) {
        SymbolicLU::Seg& g = s.split_bseg[j];
        g.prog_b_off = static_cast<int>(s.prog_bs.size());
        const auto& rv = brecs[static_cast<size_t>(j)];
        for (size_t q = 0; q < rv.size(); ++q) {
            const BRec& rec = rv[q];
            const int next = q + 1 < rv.size() ? head(rv[q + 1]) : 0;
            const size_t at = s.prog_bs.size();
            if (rec.kind == 1) {
                const int m = static_cast<int>(rec.r.size() / 2);
                s.prog_bs.resize(at + static_cast<size_t>(pair_rec_words(m)), enc(-1));
                int* wd = s.prog_bs.data() + at;
                wd[0] = kPreloadHead | (next << 16);
                wd[1] = m;
                for (int i = 0; i < m; ++i) {
                    wd[2 + 2 * i] = enc(rec.r[2 * i]);
                    wd[3 + 2 * i] = rec.r[2 * i + 1];
                }
                continue;
            }
            const int c = rec.c;
            const int nw = rec.w == 2 ? brec2_words(c) : brec_words(c);
            s.prog_bs.resize(at + static_cast<size_t>(nw), enc(-1));
            int* wd = s.prog_bs.data() + at;
            wd[0] = rec_head(c, rec.w) | (next << 16);
            if (rec.w == 2) {
                wd[1] = s.perm[rec.k + 1];
                wd[2] = s.perm[rec.k];
                for (size_t i = 0; i < rec.r.size(); ++i) wd[3 + i] = enc(rec.r[i]);
            } else {
                wd[1] = s.perm[rec.k];
                for (size_t i = 0; i < rec.r.size(); ++i) wd[2 + i] = enc(rec.r[i]);
            }
        }
        s.prog_bs.resize(static_cast<size_t>(round_up_to(static_cast<int>(s.prog_bs.size()), kProgChunk)), enc(-1));
        g.prog_b_len = static_cast<int>(s.prog_bs.size()) - g.prog_b_off;
        g.u_off = u0[j];
        g.u_len = u1[j] - u0[j];
        g.n_piv = piv[j];
    }
}

}  // namespace

namespace {

// Checks an extension spec against the grid (bus kinds: -1 fixed, 0 pv, 1 pq).
void check_ext(const ExtSpec& x, int n, int slack, const std::vector<int>& kind)
{
    auto bus_ok = [&](int b) { return b >= 0 && b < n; };
    if (x.has_ds()) {
        if (static_cast<int>(x.slack_w.size()) != n) throw std::invalid_argument("slack_weights needs n entries");
        for (int i = 0; i < n; ++i) {
            if (!std::isfinite(x.slack_w[i])) throw std::invalid_argument("slack weight is not finite");
            if (x.slack_w[i] != 0.0 && kind[i] < 0 && i != slack)
                throw std::invalid_argument("a distributed-slack participant must be a pv or pq bus or the slack");
        }
    }
    const int nh = x.n_hvdc();
    for (const auto* v : {&x.h_bus2, &x.h_status})
        if (static_cast<int>(v->size()) != nh) throw std::invalid_argument("hvdc arrays differ in length");
    for (const auto* v : {&x.h_p0, &x.h_k, &x.h_lf1, &x.h_lf2, &x.h_r, &x.h_pmax12, &x.h_pmax21})
        if (static_cast<int>(v->size()) != nh) throw std::invalid_argument("hvdc arrays differ in length");
    for (int e = 0; e < nh; ++e) {
        if (!bus_ok(x.h_bus1[e]) || !bus_ok(x.h_bus2[e]) || x.h_bus1[e] == x.h_bus2[e])
            throw std::invalid_argument("bad hvdc end bus");
        const int st = x.h_status[e];
        if (st != 0 && st != 1 && st != -1 && st != kHvdcTripped) throw std::invalid_argument("bad hvdc status");
    }
    const int ng = x.n_grp(), nc = x.n_ctrl();
    if (ng > 0) {
        if (static_cast<int>(x.vc_vset.size()) != ng || static_cast<int>(x.vc_ptr.size()) != ng + 1 ||
            x.vc_ptr[0] != 0 || x.vc_ptr[ng] != nc || static_cast<int>(x.vc_slope.size()) != nc ||
            static_cast<int>(x.vc_w.size()) != nc)
            throw std::invalid_argument("voltage-control arrays are inconsistent");
        for (int g = 0; g < ng; ++g) {
            if (x.vc_ptr[g + 1] <= x.vc_ptr[g]) throw std::invalid_argument("a voltage-control group has no controller");
            if (!bus_ok(x.vc_reg[g]) || kind[x.vc_reg[g]] != 1)
                throw std::invalid_argument("a regulated bus must be a pq bus");
            for (int g2 = 0; g2 < g; ++g2)
                if (x.vc_reg[g2] == x.vc_reg[g]) throw std::invalid_argument("two groups regulate one bus");
        }
        for (int c = 0; c < nc; ++c) {
            if (!bus_ok(x.vc_bus[c]) || kind[x.vc_bus[c]] != 1)
                throw std::invalid_argument("a voltage-control controller must sit at a pq bus");
            if (!(x.vc_w[c] > 0.0) || !std::isfinite(x.vc_slope[c]))
                throw std::invalid_argument("controller weights must be positive, slopes finite");
        }
    } else if (nc > 0) {
        throw std::invalid_argument("voltage-control controllers without groups");
    }
}

}  // namespace

NrPlan build_nr_plan(int n, const int* Yp, const int* Yi, const double* Yx, int npv, const int* pv,
                     int npq, const int* pq, int slack, int walkers, int bwalkers, const ExtSpec* ext)
{
    if (n <= 1) throw std::invalid_argument("the grid needs at least two buses");
    NrPlan P;
    P.n = n;
    P.nY = Yp[n];
    P.npv = npv;
    P.npq = npq;
    P.npvpq = npv + npq;
    P.slack = slack;
    P.Yp.assign(Yp, Yp + n + 1);
    P.Yi.assign(Yi, Yi + P.nY);
    if (Yx) P.Yx.assign(Yx, Yx + 2 * static_cast<size_t>(P.nY));
    for (int i = 0; i < n; ++i) {
        if (Yp[i + 1] < Yp[i]) throw std::invalid_argument("Ybus indptr is not monotone");
        for (int p = Yp[i]; p < Yp[i + 1]; ++p) {
            if (Yi[p] < 0 || Yi[p] >= n) throw std::invalid_argument("Ybus column index out of range");
            if (p > Yp[i] && Yi[p] <= Yi[p - 1])
                throw std::invalid_argument("Ybus columns must be sorted and unique per row");
        }
    }
    std::vector<int> kind(static_cast<size_t>(n), -1);  // 0 pv, 1 pq
    for (int k = 0; k < npv; ++k) {
        if (pv[k] < 0 || pv[k] >= n || kind[pv[k]] >= 0) throw std::invalid_argument("bad or repeated pv bus");
        kind[pv[k]] = 0;
    }
    for (int k = 0; k < npq; ++k) {
        if (pq[k] < 0 || pq[k] >= n || kind[pq[k]] >= 0) throw std::invalid_argument("bad or repeated pq bus");
        kind[pq[k]] = 1;
    }
    // every bus outside pv and pq is fixed (|V| and angle, no unknowns), as
    // the reference's array mode treats them (acpf_nr.cu:644-647,1732-1733);
    // `slack` names the fixed bus that anchors the angles
    if (slack < 0 || slack >= n || kind[slack] >= 0) throw std::invalid_argument("bad slack bus");
    if (ext && !ext->empty()) {
        check_ext(*ext, n, slack, kind);
        P.ext = *ext;
    }
    const ExtSpec& X = P.ext;
    const int nc = X.n_ctrl(), ng = X.n_grp();
    // Layout.  Rows [P of sorted pvpq | Q of pq in input order] and columns
    // [theta of pvpq | |V| of pq] (acpf_nr.cu:645-647,678-685), then the
    // extension rows / columns, paired by index (each pair is a diagonal
    // entry of the static-pivot factorization):
    //   distributed slack: the slack's P row  <->  sa (the absorbed power);
    //   per voltage-control group with controllers c_1..c_N: the voltage row
    //   <-> Q(c_1), then the sharing row of c_k <-> Q(c_k), k = 2..N.
    P.pvpq.assign(pv, pv + npv);
    P.pvpq.insert(P.pvpq.end(), pq, pq + npq);
    std::sort(P.pvpq.begin(), P.pvpq.end());
    P.pq.assign(pq, pq + npq);
    P.row_p.assign(static_cast<size_t>(n), -1);
    P.row_q.assign(static_cast<size_t>(n), -1);
    P.col_th.assign(static_cast<size_t>(n), -1);
    for (int k = 0; k < P.npvpq; ++k) P.row_p[P.pvpq[k]] = P.col_th[P.pvpq[k]] = k;
    for (int k = 0; k < npq; ++k) P.row_q[pq[k]] = P.npvpq + k;
    const std::vector<int>& col_vm = P.row_q;
    int nxt = P.npvpq + npq;
    ExtTables& xt = P.xt;
    if (X.has_ds()) {
        P.row_p[slack] = xt.sa_col = nxt++;
        xt.sa = xt.n_xs++;
    }
    std::vector<int> v_row(static_cast<size_t>(ng), -1), q_col(static_cast<size_t>(nc), -1);
    xt.q0 = xt.n_xs;
    xt.n_xs += nc;
    for (int g = 0; g < ng; ++g)
        for (int c = X.vc_ptr[g]; c < X.vc_ptr[g + 1]; ++c) {
            q_col[c] = nxt++;
            if (c == X.vc_ptr[g]) v_row[g] = q_col[c];
        }
    P.n_ext = nxt - (P.npvpq + npq);
    P.d = nxt;
    const int d = P.d;

    // J pattern: every (row, col) a coefficient lands on, with its sources
    struct Src { int col, kind, a, b; };  // kind 0: Ybus entry a, block b; 1: extension record a
    enum { kY = 0, kX = 1 };
    std::vector<std::vector<Src>> rows(static_cast<size_t>(d));
    for (int i = 0; i < n; ++i)
        for (int a = 0; a < 2; ++a) {
            const int r = a == 0 ? P.row_p[i] : P.row_q[i];
            if (r < 0) continue;
            for (int p = Yp[i]; p < Yp[i + 1]; ++p) {
                const int j = Yi[p];
                if (P.col_th[j] >= 0) rows[r].push_back({P.col_th[j], kY, p, a == 0 ? 0 : 2});
                if (col_vm[j] >= 0) rows[r].push_back({col_vm[j], kY, p, a == 0 ? 1 : 3});
            }
        }
    // extension J sources: {row, col, kind, index} (kind: kExtJConst with a
    // value, kExtJHvdc with 4 * line + q)
    struct XJ { int row, col, kind, idx; double val; };
    std::vector<XJ> xj;
    if (X.has_ds())
        for (int i = 0; i < n; ++i)
            if (X.slack_w[i] != 0.0 || i == slack) xj.push_back({P.row_p[i], xt.sa_col, kExtJConst, 0, X.slack_w[i]});
    for (int e = 0; e < X.n_hvdc(); ++e) {
        if (X.h_status[e] != 0) continue;  // saturated or tripped: constant flow, no slope
        const int bb[2] = {X.h_bus1[e], X.h_bus2[e]};
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b) {
                const int r = P.row_p[bb[a]], c = P.col_th[bb[b]];
                if (r >= 0 && c >= 0) xj.push_back({r, c, kExtJHvdc, 4 * e + 2 * a + b, 0.0});
            }
    }
    for (int g = 0; g < ng; ++g) {
        const int c1 = X.vc_ptr[g];
        xj.push_back({v_row[g], col_vm[X.vc_reg[g]], kExtJConst, 0, 1.0});
        for (int c = c1; c < X.vc_ptr[g + 1]; ++c) {
            xj.push_back({P.row_q[X.vc_bus[c]], q_col[c], kExtJConst, 0, -1.0});
            xj.push_back({v_row[g], q_col[c], kExtJConst, 0, X.vc_slope[c]});
            if (c > c1) {  // sharing row of c: w_1 Q_c - w_c Q_1 = 0
                xj.push_back({q_col[c], q_col[c], kExtJConst, 0, X.vc_w[c1]});
                xj.push_back({q_col[c], q_col[c1], kExtJConst, 0, -X.vc_w[c]});
            }
        }
    }
    for (size_t q = 0; q < xj.size(); ++q) rows[xj[q].row].push_back({xj[q].col, kX, static_cast<int>(q), 0});
    P.Jp.assign(static_cast<size_t>(d) + 1, 0);
    for (int b = 0; b < 4; ++b) P.map[b].assign(static_cast<size_t>(P.nY), -1);
    std::vector<int> xj_pos(xj.size(), -1);
    std::vector<char> pos_has_y;  // per CSR position: a Ybus source writes it
    for (int r = 0; r < d; ++r) {
        auto& v = rows[r];
        std::stable_sort(v.begin(), v.end(), [](const Src& x, const Src& y) { return x.col < y.col; });
        for (size_t q = 0; q < v.size(); ++q) {
            if (q == 0 || v[q].col != v[q - 1].col) {
                P.Ji.push_back(v[q].col);
                pos_has_y.push_back(0);
            }
            const int pos = static_cast<int>(P.Ji.size()) - 1;
            if (v[q].kind == kY) {
                P.map[v[q].b][v[q].a] = pos;
                pos_has_y[pos] = 1;
            } else {
                xj_pos[v[q].a] = pos;
            }
        }
        P.Jp[r + 1] = static_cast<int>(P.Ji.size());
    }
    P.nJ = P.Jp[d];

    // bus-level minimum degree over the buses with unknowns (P and Q of a
    // bus eliminated together), weight = unknowns per bus; the extension
    // couplings (HVDC ends, controllers and their regulated bus) are edges
    std::vector<int> bus_ids, bus_idx(static_cast<size_t>(n), -1);
    for (int i = 0; i < n; ++i)
        if (kind[i] >= 0) { bus_idx[i] = static_cast<int>(bus_ids.size()); bus_ids.push_back(i); }
    const int nb = static_cast<int>(bus_ids.size());
    std::vector<int> badjp(static_cast<size_t>(nb) + 1, 0), badji, w(static_cast<size_t>(nb));
    {
        std::vector<std::vector<int>> nbv(static_cast<size_t>(nb));
        auto edge = [&](int i, int j) {
            if (i == j || bus_idx[i] < 0 || bus_idx[j] < 0) return;
            nbv[bus_idx[i]].push_back(bus_idx[j]);
            nbv[bus_idx[j]].push_back(bus_idx[i]);
        };
        for (int b = 0; b < nb; ++b) {
            const int i = bus_ids[b];
            w[b] = kind[i] == 1 ? 2 : 1;
            for (int p = Yp[i]; p < Yp[i + 1]; ++p) edge(i, Yi[p]);
        }
        for (int e = 0; e < X.n_hvdc(); ++e) edge(X.h_bus1[e], X.h_bus2[e]);
        for (int g = 0; g < ng; ++g)
            for (int c = X.vc_ptr[g]; c < X.vc_ptr[g + 1]; ++c) {
                edge(X.vc_bus[c], X.vc_reg[g]);
                edge(X.vc_bus[c], X.vc_bus[X.vc_ptr[g]]);
            }
        for (int b = 0; b < nb; ++b) {
            auto& v = nbv[b];
            std::sort(v.begin(), v.end());
            v.erase(std::unique(v.begin(), v.end()), v.end());
            badji.insert(badji.end(), v.begin(), v.end());
            badjp[b + 1] = static_cast<int>(badji.size());
        }
    }
    std::vector<int> border = min_degree_order(nb, badjp, badji, w);
    std::vector<int> order;
    order.reserve(static_cast<size_t>(d));
    for (int b : border) {
        const int i = bus_ids[b];
        order.push_back(P.row_p[i]);
        if (P.row_q[i] >= 0) order.push_back(P.row_q[i]);
    }
    // extension pivots last: each group's sharing rows, then its voltage row
    // (its pivot is the group's weighted sensitivity once the buses and the
    // sharing rows are eliminated), the distributed slack at the very end
    for (int g = 0; g < ng; ++g) {
        for (int c = X.vc_ptr[g] + 1; c < X.vc_ptr[g + 1]; ++c) order.push_back(q_col[c]);
        order.push_back(v_row[g]);
    }
    if (X.has_ds()) order.push_back(xt.sa_col);
    if (static_cast<int>(order.size()) != d) throw std::logic_error("plan: the pivot order misses unknowns");
    P.lu = build_symbolic(d, P.Jp.data(), P.Ji.data(), order, walkers > 0 ? walkers : default_walkers(d),
                          bwalkers > 0 ? bwalkers : default_bwalkers(d));

    // extension tables in A-stream positions
    if (!X.empty()) {
        std::vector<int> xb_of(static_cast<size_t>(n), -1);
        std::vector<std::vector<std::pair<int, int>>> inj;
        auto add_inj = [&](int bus, int k, int idx) {
            if (xb_of[bus] < 0) {
                xb_of[bus] = static_cast<int>(xt.xb_bus.size());
                xt.xb_bus.push_back(bus);
                inj.emplace_back();
            }
            inj[xb_of[bus]].push_back({k, idx});
        };
        if (X.has_ds())
            for (int i = 0; i < n; ++i)
                if (X.slack_w[i] != 0.0) {
                    add_inj(i, kExtInjDs, static_cast<int>(xt.ds_w.size()));
                    xt.ds_w.push_back(X.slack_w[i]);
                }
        for (int e = 0; e < X.n_hvdc(); ++e) {
            if (P.row_p[X.h_bus1[e]] >= 0) add_inj(X.h_bus1[e], kExtInjHvdc1, e);
            if (P.row_p[X.h_bus2[e]] >= 0) add_inj(X.h_bus2[e], kExtInjHvdc2, e);
        }
        for (int c = 0; c < nc; ++c) add_inj(X.vc_bus[c], kExtInjVc, c);
        xt.inj_ptr.assign(1, 0);
        for (const auto& v : inj) {
            for (const auto& [k, idx] : v) xt.inj.insert(xt.inj.end(), {k, idx});
            xt.inj_ptr.push_back(static_cast<int>(xt.inj.size() / 2));
        }
        // J records, the first of a position without a Ybus source stores
        std::vector<char> seen(static_cast<size_t>(P.nJ), 0);
        for (size_t q = 0; q < xj.size(); ++q) {
            const int pos = xj_pos[q];
            const int apos = P.lu.a_pos_of_entry[pos];
            if (pos < 0 || apos < 0 || apos >= P.lu.a_len) throw std::logic_error("plan: extension entry off the A stream");
            const int add = (pos_has_y[pos] || seen[pos]) ? 1 : 0;
            seen[pos] = 1;
            int idx = xj[q].idx;
            if (xj[q].kind == kExtJConst) {
                idx = static_cast<int>(xt.jconst.size());
                xt.jconst.push_back(xj[q].val);
            }
            xt.jrec.insert(xt.jrec.end(), {apos, xj[q].kind, idx, add});
        }
        for (int r = P.npvpq + npq; r < d; ++r)
            if (P.lu.f_pos[r] < 0 || P.lu.f_pos[r] >= P.lu.a_len) throw std::logic_error("plan: extension row off the A stream");
        for (int g = 0; g < ng; ++g) {
            xt.frec.insert(xt.frec.end(), {P.lu.f_pos[v_row[g]], kExtFVc, g});
            for (int c = X.vc_ptr[g] + 1; c < X.vc_ptr[g + 1]; ++c) {
                xt.frec.insert(xt.frec.end(), {P.lu.f_pos[q_col[c]], kExtFShare, static_cast<int>(xt.sh_c1.size())});
                xt.sh_c1.push_back(X.vc_ptr[g]);
                xt.sh_ck.push_back(c);
            }
        }
        xt.xcol.assign(static_cast<size_t>(xt.n_xs), -1);
        if (xt.sa >= 0) xt.xcol[xt.sa] = xt.sa_col;
        for (int c = 0; c < nc; ++c) xt.xcol[xt.q0 + c] = q_col[c];
        xt.h_bus1 = X.h_bus1;
        xt.h_bus2 = X.h_bus2;
        xt.h_status = X.h_status;
        for (int e = 0; e < X.n_hvdc(); ++e)
            xt.h_par.insert(xt.h_par.end(), {X.h_p0[e], X.h_k[e], X.h_lf1[e], X.h_lf2[e], X.h_r[e], X.h_pmax12[e],
                                             X.h_pmax21[e]});
        xt.vc_reg = X.vc_reg;
        xt.vc_ptr = X.vc_ptr;
        xt.vc_vset = X.vc_vset;
        xt.vc_slope = X.vc_slope;
        xt.vc_w = X.vc_w;
    }

    // mismatch/Jacobian program: buses with a P equation in RCM order of the Ybus graph
    {
        std::vector<int> xb_of(static_cast<size_t>(n), -1);
        for (size_t q = 0; q < xt.xb_bus.size(); ++q) xb_of[xt.xb_bus[q]] = static_cast<int>(q);
        std::vector<int> yadjp(static_cast<size_t>(n) + 1, 0), yadji;
        for (int i = 0; i < n; ++i) {
            for (int p = Yp[i]; p < Yp[i + 1]; ++p)
                if (Yi[p] != i) yadji.push_back(Yi[p]);
            yadjp[i + 1] = static_cast<int>(yadji.size());
        }
        std::vector<int> ro = rcm_order(n, yadjp, yadji);
        // the tables for one stream layout (A-stream positions of the J
        // entries and of the right-hand sides)
        auto k1_tables = [&](const std::vector<int>& apos, const std::vector<int>& fpos, std::vector<int>& kb,
                             std::vector<int>& ke) {
            kb.clear();
            ke.clear();
            for (int i : ro) {
                if (P.row_p[i] < 0) continue;  // fixed bus: no equations
                const int e0 = static_cast<int>(ke.size() / 6);
                int diag = -1;
                for (int p = Yp[i]; p < Yp[i + 1]; ++p) {
                    const int j = Yi[p];
                    if (j == i) diag = p;
                    int a[4];
                    for (int b = 0; b < 4; ++b) a[b] = P.map[b][p] >= 0 ? apos[P.map[b][p]] : -1;
                    ke.insert(ke.end(), {j, p, a[0], a[1], a[2], a[3]});
                }
                if (diag < 0) throw std::invalid_argument("Ybus has no diagonal entry for a bus with equations");
                const int e1 = static_cast<int>(ke.size() / 6);
                const int fp = fpos[P.row_p[i]];
                const int fq = P.row_q[i] >= 0 ? fpos[P.row_q[i]] : -1;
                kb.insert(kb.end(), {i, e0, e1, fp, fq, diag, xb_of[i], Yp[i]});  // its entries: Ybus row i
            }
        };
        k1_tables(P.lu.a_pos_of_entry, P.lu.f_pos, P.k1_bus, P.k1_ent);
        P.k1_nbus = static_cast<int>(P.k1_bus.size() / kK1BusWords);
        // the split walk's layout (plans without extension rows)
        if (P.lu.n_split > 0 && X.empty()) k1_tables(P.lu.split_apos, P.lu.split_fpos, P.k1_bus_s, P.k1_ent_s);
    }
    return P;
}

}  // namespace g2g
