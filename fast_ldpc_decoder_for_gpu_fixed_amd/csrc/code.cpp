// code.cpp -- runtime code tables, seeded generators, colouring and the
// level planner (host C++).
#include "code.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <queue>

namespace ldpc {

namespace {

// splitmix64: small, fast, and identical on every platform.
struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    // unbiased integer in [0, n)
    uint32_t below(uint32_t n) {
        if (n <= 1) return 0;
        const uint64_t lim = (~0ull) - ((~0ull) % n);
        uint64_t x;
        do x = next(); while (x >= lim);
        return (uint32_t)(x % n);
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
};


// ---------------------------------------------------------------------------
// Colour-count reduction for the check conflict graph (nbr[c]: the checks
// sharing a variable with c; colour[c] in [0, k)).
// ---------------------------------------------------------------------------
// Iterated greedy (Culberson and Luo, 1996): first-fit over the checks listed
// class by class never needs more colours than there are classes (a class is
// independent, so its checks fit the colours its first check could take), and
// listing the classes in reverse or by size lets classes merge.  Returns the
// colour count.
int iterated_greedy(const std::vector<std::vector<int>> &nbr, std::vector<int> &colour, int k, int rounds) {
    const int C = (int)colour.size();
    std::vector<int> order((size_t)C), mark((size_t)k + 1, -1);
    int stall = 0;
    for (int r = 0; r < rounds; r++) {
        std::vector<int> size((size_t)k, 0);
        for (int c : colour) size[c]++;
        std::vector<int> cls((size_t)k);
        std::iota(cls.begin(), cls.end(), 0);
        if (r % 2 == 0)
            std::reverse(cls.begin(), cls.end());
        else
            std::stable_sort(cls.begin(), cls.end(), [&](int a, int b) { return size[a] > size[b]; });
        std::vector<int> rank((size_t)k);
        for (int i = 0; i < k; i++) rank[cls[i]] = i;
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return rank[colour[a]] < rank[colour[b]]; });
        std::vector<int> next((size_t)C, -1);
        int kn = 0;
        for (int c : order) {
            for (int b : nbr[c])
                if (next[b] >= 0) mark[next[b]] = c;
            int x = 0;
            while (mark[x] == c) x++;
            next[c] = x;
            kn = std::max(kn, x + 1);
        }
        colour.swap(next);
        stall = kn < k ? 0 : stall + 1;
        k = kn;
        if (stall == 4) break;  // four rounds without a merge
        mark.assign((size_t)k + 1, -1);
    }
    return k;
}

// TabuCol (Hertz and de Werra, 1987) for a proper colouring with k colours,
// from a colouring in [0, k) that may have conflicts: each move recolours a
// conflicting check to the colour with the fewest conflicts (the best move
// over all conflicting checks, ties broken at random), and forbids that check
// to return to the colour it left for 10 + 0.6 * (conflicting checks) moves,
// unless the move would remove every conflict.  True when a proper colouring
// is found within max_moves.
bool tabucol(const std::vector<std::vector<int>> &nbr, std::vector<int> &colour, int k, long max_moves,
             Rng &rng) {
    const int C = (int)colour.size();
    if (k < 2) {  // one colour: no move exists; proper only with no edge at all
        for (int c = 0; c < C; c++)
            if (!nbr[c].empty()) return false;
        return true;
    }
    std::vector<int> gamma((size_t)C * k, 0);  // gamma[c*k + x]: neighbours of c with colour x
    for (int c = 0; c < C; c++)
        for (int b : nbr[c]) gamma[(size_t)c * k + colour[b]]++;
    std::vector<long> tabu((size_t)C * k, 0);
    std::vector<int> conf, where((size_t)C, -1);
    auto update = [&](int c) {
        const bool bad = gamma[(size_t)c * k + colour[c]] > 0;
        if (bad && where[c] < 0) {
            where[c] = (int)conf.size();
            conf.push_back(c);
        } else if (!bad && where[c] >= 0) {
            const int last = conf.back();
            conf[where[c]] = last;
            where[last] = where[c];
            conf.pop_back();
            where[c] = -1;
        }
    };
    long edges_in_conflict = 0;
    for (int c = 0; c < C; c++) {
        update(c);
        edges_in_conflict += gamma[(size_t)c * k + colour[c]];
    }
    edges_in_conflict /= 2;
    for (long it = 0; it < max_moves; it++) {
        if (conf.empty()) return true;
        int bv = -1, bx = -1, bd = 1 << 30, ties = 0;
        for (int c : conf) {
            const int *g = &gamma[(size_t)c * k];
            const int own = g[colour[c]];
            for (int x = 0; x < k; x++) {
                if (x == colour[c]) continue;
                const int d = g[x] - own;
                if (tabu[(size_t)c * k + x] > it && edges_in_conflict + d > 0) continue;
                if (d < bd) {
                    bd = d, bv = c, bx = x, ties = 1;
                } else if (d == bd && rng.below((uint32_t)++ties) == 0) {
                    bv = c, bx = x;
                }
            }
        }
        if (bv < 0) {  // every move tabu: a random one
            bv = conf[rng.below((uint32_t)conf.size())];
            bx = (colour[bv] + 1 + (int)rng.below((uint32_t)k - 1)) % k;
            bd = gamma[(size_t)bv * k + bx] - gamma[(size_t)bv * k + colour[bv]];
        }
        const int old = colour[bv];
        colour[bv] = bx;
        edges_in_conflict += bd;
        for (int b : nbr[bv]) {
            gamma[(size_t)b * k + old]--;
            gamma[(size_t)b * k + bx]++;
            update(b);
        }
        update(bv);
        tabu[(size_t)bv * k + old] = it + 10 + (long)rng.below(10) + (long)(0.6 * (double)conf.size());
    }
    return conf.empty();
}

// Even class sizes: a check of a class above ceil(C / k) moves to the smallest
// class none of its neighbours has, while that class is below the bound.
void balance_classes(const std::vector<std::vector<int>> &nbr, std::vector<int> &colour, int k) {
    const int C = (int)colour.size();
    const int bound = (C + k - 1) / k;
    std::vector<int> size((size_t)k, 0);
    for (int c : colour) size[c]++;
    std::vector<char> seen((size_t)k);
    for (int pass = 0; pass < 4; pass++) {
        bool moved = false;
        for (int c = 0; c < C; c++) {
            if (size[colour[c]] <= bound) continue;
            std::fill(seen.begin(), seen.end(), 0);
            for (int b : nbr[c]) seen[colour[b]] = 1;
            int best = -1;
            for (int x = 0; x < k; x++)
                if (!seen[x] && size[x] < bound && (best < 0 || size[x] < size[best])) best = x;
            if (best < 0) continue;
            size[colour[c]]--;
            colour[c] = best;
            size[best]++;
            moved = true;
        }
        if (!moved) break;
    }
}

}  // namespace

std::string finish_code(ldpc_code &c, int q1_split) {
    if (c.N <= 0 || c.C <= 0) return "code needs at least one variable and one check";
    if ((int)c.deg.size() != c.C) return "degree list length != number of checks";
    c.eoff.assign((size_t)c.C + 1, 0);
    c.max_deg = 0;
    for (int i = 0; i < c.C; i++) {
        if (c.deg[i] < 1) return "check degree must be >= 1";
        c.eoff[i + 1] = c.eoff[i] + (size_t)c.deg[i];
        c.max_deg = std::max(c.max_deg, c.deg[i]);
    }
    if (c.eoff[c.C] != c.pos.size()) return "edge list length != sum of check degrees";
    if (c.pos.size() > 0x7FFFFFFFull) return "too many edges";
    c.E = (int)c.pos.size();
    for (uint32_t v : c.pos)
        if (v >= (uint32_t)c.N) return "variable index out of range";
    if (q1_split < 0) {
        int i = 1;
        while (i < c.C && c.deg[i] == c.deg[0]) i++;
        q1_split = i;  // == C for a single-degree code
    }
    c.q1_split = std::min(q1_split, c.C);
    return "";
}

int dependency_levels(const ldpc_code &c, std::vector<int> *level_of) {
    std::vector<int> last((size_t)c.N, -1);
    if (level_of) level_of->assign((size_t)c.C, 0);
    int n_levels = 0;
    for (int i = 0; i < c.C; i++) {
        int lvl = 0;
        for (size_t e = c.eoff[i]; e < c.eoff[i + 1]; e++) lvl = std::max(lvl, last[c.pos[e]] + 1);
        for (size_t e = c.eoff[i]; e < c.eoff[i + 1]; e++) last[c.pos[e]] = lvl;
        if (level_of) (*level_of)[i] = lvl;
        n_levels = std::max(n_levels, lvl + 1);
    }
    return n_levels;
}

// Level sizes rounded to whole granules of checks, the table's order kept.
// Two checks that share a variable keep their table order (the earlier one
// sits on a lower level), so every assignment built here decodes bit for bit
// like the serial schedule; only checks with slack move.  List scheduling
// over the n_levels of `dependency_levels`: a check is ready once every
// earlier check sharing a variable with it sits on a lower level; its latest
// level (alap) is one below the earliest latest level of the later checks
// sharing a variable with it.  Level l takes the ready checks whose latest
// level is l, then ready checks by (alap, table index) up to the multiple of
// `granule` nearest to the remaining checks per remaining level (all that is
// ready on the last level).  Placing every check between its ready level and
// its alap keeps the number of levels.
static void balance_levels(const ldpc_code &c, int n_levels, int granule, std::vector<int> &lvl) {
    const int C = c.C;
    std::vector<std::vector<int>> succ((size_t)C);
    std::vector<int> npred((size_t)C, 0), last((size_t)c.N, -1);
    for (int i = 0; i < C; i++) {
        std::vector<int> pr;
        for (size_t e = c.eoff[i]; e < c.eoff[i + 1]; e++) {
            const int q = last[c.pos[e]];
            if (q >= 0) pr.push_back(q);
            last[c.pos[e]] = i;
        }
        std::sort(pr.begin(), pr.end());
        pr.erase(std::unique(pr.begin(), pr.end()), pr.end());
        npred[i] = (int)pr.size();
        for (int q : pr) succ[(size_t)q].push_back(i);
    }
    std::vector<int> alap((size_t)C, n_levels - 1);
    for (int i = C - 1; i >= 0; i--)
        for (int s : succ[(size_t)i]) alap[i] = std::min(alap[i], alap[s] - 1);
    std::vector<int> ready;
    for (int i = 0; i < C; i++)
        if (npred[i] == 0) ready.push_back(i);
    int left = C;
    for (int l = 0; l < n_levels; l++) {
        std::sort(ready.begin(), ready.end(), [&](int a, int b) {
            return alap[a] != alap[b] ? alap[a] < alap[b] : a < b;
        });
        const int r = (int)ready.size();
        int forced = 0;
        while (forced < r && alap[ready[forced]] <= l) forced++;
        int n = r;
        if (l < n_levels - 1) {
            const double target = (double)left / (n_levels - l);
            const int lo = (forced + granule - 1) / granule * granule;
            if (lo <= r) {
                n = lo;
                for (int m = lo; m <= r; m += granule)
                    if (std::abs(m - target) < std::abs(n - target)) n = m;
            }
        }
        std::vector<int> next(ready.begin() + n, ready.end());
        for (int k = 0; k < n; k++) lvl[ready[k]] = l;
        for (int k = 0; k < n; k++)
            for (int s : succ[(size_t)ready[k]])
                if (--npred[s] == 0) next.push_back(s);
        left -= n;
        ready.swap(next);
    }
}

Plan make_plan(const ldpc_code &c, int granule) {
    Plan p;
    std::vector<int> lvl;
    p.n_levels = dependency_levels(c, &lvl);
    if (granule > 1 && p.n_levels > 1) balance_levels(c, p.n_levels, granule, lvl);
    p.max_deg = c.max_deg;
    p.order.resize((size_t)c.C);
    std::iota(p.order.begin(), p.order.end(), 0);
    // by level, then degree descending (keeps a wavefront on one degree), then table index
    std::stable_sort(p.order.begin(), p.order.end(), [&](int a, int b) {
        if (lvl[a] != lvl[b]) return lvl[a] < lvl[b];
        return c.deg[a] > c.deg[b];
    });
    p.level_beg.assign((size_t)p.n_levels + 1, 0);
    for (int i = 0; i < c.C; i++) p.level_beg[lvl[i] + 1]++;
    for (int l = 0; l < p.n_levels; l++) p.level_beg[l + 1] += p.level_beg[l];
    p.level_deg.assign((size_t)p.n_levels, -1);
    for (int i = 0; i < c.C; i++) {
        int &d = p.level_deg[lvl[i]];
        d = d < 0 ? c.deg[i] : (d == c.deg[i] ? d : 0);
    }
    p.slot_of.resize((size_t)c.C);
    p.check_info.resize((size_t)c.C);
    p.check_e0.resize((size_t)c.C);
    p.edge_var.resize((size_t)c.E);
    uint32_t e0 = 0;
    for (int s = 0; s < c.C; s++) {
        const int t = p.order[s];
        p.slot_of[t] = s;
        p.check_info[s] = (uint32_t)c.deg[t] | ((t >= c.q1_split) ? 0x100u : 0u);
        p.check_e0[s] = e0;
        for (int j = 0; j < c.deg[t]; j++) p.edge_var[e0 + j] = c.pos[c.eoff[t] + j];
        e0 += (uint32_t)c.deg[t];
    }
    return p;
}

// ---------------------------------------------------------------------------
// regular (dv, dc) code: check sockets filled one by one from the pool of
// variable sockets (a random permutation drawn lazily), rejecting a variable
// already in the check or one that would close a 4-cycle.
// ---------------------------------------------------------------------------
std::string gen_regular(ldpc_code &out, int n, int dv, int dc, uint64_t seed) {
    if (n <= 0 || dv <= 0 || dc <= 1) return "bad regular code parameters";
    if (((long long)n * dv) % dc) return "n*dv must be a multiple of dc";
    const int C = (int)((long long)n * dv / dc);
    for (int attempt = 0; attempt < 64; attempt++) {
        Rng rng(seed * 0x9E3779B97F4A7C15ull + (uint64_t)attempt * 0xD1B54A32D192ED03ull + 1);
        std::vector<uint32_t> pool((size_t)n * dv);
        for (int v = 0; v < n; v++)
            for (int k = 0; k < dv; k++) pool[(size_t)v * dv + k] = (uint32_t)v;
        size_t remaining = pool.size();
        std::vector<std::vector<int>> var_checks((size_t)n);
        std::vector<uint32_t> pos;
        pos.reserve(pool.size());
        bool ok = true;
        std::vector<char> mark((size_t)n, 0);  // variables sharing a check with the current check's vars
        for (int c = 0; c < C && ok; c++) {
            const size_t base = pos.size();
            for (int j = 0; j < dc && ok; j++) {
                // variables forbidden for this slot: in the check, or 2 hops away
                auto allowed = [&](uint32_t v) {
                    for (size_t e = base; e < pos.size(); e++)
                        if (pos[e] == v) return false;
                    return !mark[v];
                };
                size_t pick = remaining;
                for (int tries = 0; tries < 64 && pick == remaining; tries++) {
                    const size_t r = rng.below((uint32_t)remaining);
                    if (allowed(pool[r])) pick = r;
                }
                if (pick == remaining) {
                    const size_t start = rng.below((uint32_t)remaining);
                    for (size_t k = 0; k < remaining; k++) {
                        const size_t r = (start + k) % remaining;
                        if (allowed(pool[r])) { pick = r; break; }
                    }
                }
                if (pick == remaining) { ok = false; break; }
                const uint32_t v = pool[pick];
                pool[pick] = pool[remaining - 1];
                remaining--;
                pos.push_back(v);
                // every variable sharing an earlier check with v is now 2 hops away
                for (int oc : var_checks[v]) {
                    const size_t b0 = (size_t)oc * dc;
                    for (int k = 0; k < dc; k++) mark[pos[b0 + k]] = 1;
                }
            }
            if (!ok) break;
            for (int j = 0; j < dc; j++) var_checks[pos[base + j]].push_back(c);
            // clear marks of this check's neighbourhood
            for (int j = 0; j < dc; j++)
                for (int oc : var_checks[pos[base + j]]) {
                    const size_t b0 = (size_t)oc * dc;
                    for (int k = 0; k < dc && b0 + k < pos.size(); k++) mark[pos[b0 + k]] = 0;
                }
            std::sort(pos.begin() + (long)base, pos.end());
        }
        if (!ok) continue;
        out = ldpc_code();
        out.N = n;
        out.C = C;
        out.deg.assign((size_t)C, dc);
        out.pos = std::move(pos);
        return finish_code(out, -1);
    }
    return "could not build a 4-cycle-free regular code (64 attempts)";
}

// ---------------------------------------------------------------------------
// quasi-cyclic code from a random base graph with a dual-diagonal parity part
// ---------------------------------------------------------------------------
std::string gen_qc(ldpc_code &out, int mb, int nb, int z, double p, uint64_t seed) {
    if (mb < 2 || nb <= mb || z < 1 || p < 0 || p > 1) return "bad QC parameters";
    Rng rng(seed ^ 0x51C0DE5EEDull);
    const int kb = nb - mb;
    std::vector<int> base((size_t)mb * nb, -1);
    // info columns: each base entry present with probability p, a column
    // redrawn until it has at least two entries (a weight-1 variable only
    // ever sees its channel value as its extrinsic input, so it can neither
    // be corrected by its other checks nor let the parity stop test pass
    // when its channel sign is wrong)
    for (int c = 0; c < kb; c++) {
        int cnt = 0;
        for (int attempt = 0; attempt < 1000 && cnt < 2; attempt++) {
            cnt = 0;
            for (int r = 0; r < mb; r++) {
                const bool present = rng.uniform() < p;
                const int shift = (int)rng.below((uint32_t)z);
                base[(size_t)r * nb + c] = present ? shift : -1;
                cnt += present;
            }
        }
        for (int r = 0; cnt < 2 && r < mb; r++)
            if (base[(size_t)r * nb + c] < 0) {
                base[(size_t)r * nb + c] = (int)rng.below((uint32_t)z);
                cnt++;
            }
    }
    // parity part: first parity column has shift 1 at the top and bottom and
    // 0 in the middle; the others form a dual diagonal of identities.
    base[(size_t)0 * nb + kb] = 1 % z;
    base[(size_t)(mb / 2) * nb + kb] = 0;
    base[(size_t)(mb - 1) * nb + kb] = 1 % z;
    for (int i = 0; i + 1 < mb; i++) {
        base[(size_t)i * nb + kb + 1 + i] = 0;
        base[(size_t)(i + 1) * nb + kb + 1 + i] = 0;
    }
    out = ldpc_code();
    out.N = nb * z;
    out.C = mb * z;
    for (int r = 0; r < mb; r++)
        for (int i = 0; i < z; i++) {
            int d = 0;
            for (int c = 0; c < nb; c++) {
                const int s = base[(size_t)r * nb + c];
                if (s < 0) continue;
                out.pos.push_back((uint32_t)(c * z + (i + s) % z));
                d++;
            }
            out.deg.push_back(d);
        }
    return finish_code(out, -1);
}

// ---------------------------------------------------------------------------
// IRA staircase code
// ---------------------------------------------------------------------------
std::string gen_ira(ldpc_code &out, int k, int n_hi, int d_hi, int d_lo, uint64_t seed) {
    if (k < 2 || n_hi < 0 || n_hi > k || d_hi < 1 || d_lo < 1) return "bad IRA parameters";
    const long long info_edges = (long long)n_hi * d_hi + (long long)(k - n_hi) * d_lo;
    if (info_edges % k) return "info edges must be a multiple of the number of checks";
    const int dci = (int)(info_edges / k);
    if (dci < 1) return "check info degree must be >= 1";
    std::vector<uint32_t> sock;
    sock.reserve((size_t)info_edges);
    for (int v = 0; v < k; v++) {
        const int d = v < n_hi ? d_hi : d_lo;
        for (int j = 0; j < d; j++) sock.push_back((uint32_t)v);
    }
    Rng rng(seed ^ 0x1EA5EEDull);
    for (size_t i = sock.size(); i > 1; i--) std::swap(sock[i - 1], sock[rng.below((uint32_t)i)]);
    // repair repeated variables inside a check by swapping with a random slot
    for (int pass = 0; pass < 100; pass++) {
        bool clean = true;
        for (int c = 0; c < k; c++) {
            uint32_t *s = &sock[(size_t)c * dci];
            for (int a = 0; a < dci; a++)
                for (int b = a + 1; b < dci; b++)
                    if (s[a] == s[b]) {
                        clean = false;
                        std::swap(s[b], sock[rng.below((uint32_t)sock.size())]);
                    }
        }
        if (clean) break;
        if (pass == 99) return "could not remove repeated variables";
    }
    out = ldpc_code();
    out.N = 2 * k;
    out.C = k;
    for (int c = 0; c < k; c++) {
        std::vector<uint32_t> vs(sock.begin() + (long)c * dci, sock.begin() + (long)(c + 1) * dci);
        std::sort(vs.begin(), vs.end());
        if (c > 0) vs.push_back((uint32_t)(k + c - 1));
        vs.push_back((uint32_t)(k + c));
        out.deg.push_back((int)vs.size());
        out.pos.insert(out.pos.end(), vs.begin(), vs.end());
    }
    return finish_code(out, -1);
}

// ---------------------------------------------------------------------------
// balanced colouring of the conflict graph (checks that share a variable
// conflict); table rewritten colour by colour.  DSatur order (most colours
// among coloured neighbours first, then most neighbours, then lowest index);
// each check takes the least-filled colour its neighbours leave free, and a
// new colour only when none is free.  It starts with as many colours as the
// largest variable degree (the checks of one variable form a clique, so no
// colouring needs fewer).  The colour count is then cut by iterated greedy and
// tabu search, and the classes are evened out.  Few, balanced classes make the
// dependency levels of the layered schedule few and evenly filled: (3,6)
// N = 4000: 6 levels (first-fit greedy colouring 9, DSatur alone 7); the IRA
// N = 64800 code: 12 (greedy 17, DSatur 13); QC 12x24 Z = 81: 8 (DSatur 12).
// ---------------------------------------------------------------------------
std::string reorder_by_colour(const ldpc_code &in, ldpc_code &out) {
    std::vector<std::vector<int>> var_checks((size_t)in.N);
    for (int c = 0; c < in.C; c++)
        for (size_t e = in.eoff[c]; e < in.eoff[c + 1]; e++) var_checks[in.pos[e]].push_back(c);
    std::vector<std::vector<int>> nbr((size_t)in.C);
    {
        std::vector<int> mark((size_t)in.C, -1);
        for (int c = 0; c < in.C; c++) {
            mark[c] = c;
            for (size_t e = in.eoff[c]; e < in.eoff[c + 1]; e++)
                for (int o : var_checks[in.pos[e]])
                    if (mark[o] != c) {
                        mark[o] = c;
                        nbr[c].push_back(o);
                    }
        }
    }
    size_t k0 = 1;
    for (const auto &v : var_checks) k0 = std::max(k0, v.size());
    std::vector<int> size(k0, 0), colour((size_t)in.C, -1);
    std::vector<std::vector<char>> used((size_t)in.C);  // used[c][k]: a neighbour of c has colour k
    std::vector<int> sat((size_t)in.C, 0);
    // max-heap on (saturation, neighbour count, -index); stale entries skipped
    struct Item {
        int sat, deg, c;
        bool operator<(const Item &o) const {
            if (sat != o.sat) return sat < o.sat;
            if (deg != o.deg) return deg < o.deg;
            return c > o.c;
        }
    };
    std::priority_queue<Item> heap;
    for (int c = 0; c < in.C; c++) heap.push({0, (int)nbr[c].size(), c});
    while (!heap.empty()) {
        const Item it = heap.top();
        heap.pop();
        const int c = it.c;
        if (colour[c] >= 0 || it.sat != sat[c]) continue;
        int best = -1;
        for (int k = 0; k < (int)size.size(); k++) {
            const bool taken = k < (int)used[c].size() && used[c][k];
            if (!taken && (best < 0 || size[k] < size[best])) best = k;
        }
        if (best < 0) {
            best = (int)size.size();
            size.push_back(0);
        }
        colour[c] = best;
        size[best]++;
        for (int b : nbr[c]) {
            if (colour[b] >= 0) continue;
            if ((int)used[b].size() <= best) used[b].resize((size_t)best + 1, 0);
            if (!used[b][best]) {
                used[b][best] = 1;
                sat[b]++;
                heap.push({sat[b], (int)nbr[b].size(), b});
            }
        }
    }
    // fewer colours: iterated greedy, then tabu search one colour at a time
    // down to the clique bound (the smallest class is dissolved into the
    // colours fewest of its checks' neighbours have); classes evened out last
    int k = iterated_greedy(nbr, colour, (int)size.size(), 64);
    Rng rng(0xC0104EDull ^ ((uint64_t)in.C << 20) ^ (uint64_t)in.E);
    const long budget = 20000 + 4L * in.C;
    // the tabu stage keeps two C x k tables (12 bytes per entry): tables
    // beyond 2^22 entries (~50 MB) keep the iterated-greedy colouring
    while (k > (int)k0 && (size_t)in.C * (size_t)k <= ((size_t)1 << 22)) {
        std::vector<int> cnt((size_t)k, 0);
        for (int c : colour) cnt[c]++;
        const int drop = (int)(std::min_element(cnt.begin(), cnt.end()) - cnt.begin());
        std::vector<int> trial = colour;
        for (int &x : trial)
            if (x == k - 1) x = drop;
            else if (x == drop) x = k - 1;
        for (int c = 0; c < in.C; c++) {
            if (trial[c] != k - 1) continue;
            std::fill(cnt.begin(), cnt.end(), 0);
            for (int b : nbr[c])
                if (trial[b] < k - 1) cnt[trial[b]]++;
            trial[c] = (int)(std::min_element(cnt.begin(), cnt.begin() + (k - 1)) - cnt.begin());
        }
        if (!tabucol(nbr, trial, k - 1, budget, rng)) break;
        colour.swap(trial);
        k--;
    }
    balance_classes(nbr, colour, k);
    std::vector<int> order((size_t)in.C);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return colour[a] < colour[b]; });
    out = ldpc_code();
    out.N = in.N;
    out.C = in.C;
    for (int c : order) {
        out.deg.push_back(in.deg[c]);
        out.pos.insert(out.pos.end(), in.pos.begin() + (long)in.eoff[c], in.pos.begin() + (long)in.eoff[c + 1]);
    }
    return finish_code(out, -1);
}

// ---------------------------------------------------------------------------
// reference table text: "/* msg = i, deg = d */  v0, v1, ..., " repeated
// (src/matrix/4000x2000/constantes_decoder.h:3-)
// ---------------------------------------------------------------------------
std::string load_table(ldpc_code &out, const char *path, int n_vars) {
    FILE *f = std::fopen(path, "rb");
    if (!f) return std::string("cannot open ") + path;
    std::string text;
    char buf[1 << 16];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, got);
    std::fclose(f);
    out = ldpc_code();
    size_t i = 0;
    const size_t n = text.size();
    uint32_t vmax = 0;
    while (true) {
        const size_t k = text.find("deg", i);
        if (k == std::string::npos) break;
        size_t q = text.find('=', k);
        if (q == std::string::npos) break;
        const int d = std::atoi(text.c_str() + q + 1);
        size_t end = text.find("*/", q);
        if (end == std::string::npos || d < 1) return "malformed table near offset " + std::to_string(k);
        i = end + 2;
        for (int j = 0; j < d; j++) {
            while (i < n && !(text[i] >= '0' && text[i] <= '9')) {
                if (text[i] == '/' && i + 1 < n && text[i + 1] == '*') return "short check in table";
                i++;
            }
            if (i >= n) return "table ends inside a check";
            const uint32_t v = (uint32_t)std::strtoul(text.c_str() + i, nullptr, 10);
            while (i < n && text[i] >= '0' && text[i] <= '9') i++;
            out.pos.push_back(v);
            vmax = std::max(vmax, v);
        }
        out.deg.push_back(d);
    }
    if (out.deg.empty()) return "no checks found in table";
    out.C = (int)out.deg.size();
    out.N = n_vars > 0 ? n_vars : (int)vmax + 1;
    return finish_code(out, -1);
}

}  // namespace ldpc
