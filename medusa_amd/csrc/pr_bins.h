// PageRank work units shared by the device kernels and the host planner of the binned
// message exchange.  Plain C++ (no device code), so the planner can be checked on its own.
//
// Reference: MPG/UserOperators.h:29-163 scatters one message per out-edge into the slot
// msgDstID[e] of a message array, then MPG/MsgCombiner.h:68-117 sums every destination's
// segment with a CUDPP segmented scan.  The binned variant keeps the message array but lays
// it out for locality: it is chunk-major (a chunk = a range of sources whose contributions
// fit one workgroup's LDS), and inside a chunk sorted by the destination's bin (a bin = a
// range of destinations whose messages fit one workgroup's LDS).  A send workgroup writes a
// contiguous tile of its chunk's messages; a combine workgroup reads its bin's run from every
// chunk (runs padded to 4 messages, so every access is 16-byte aligned).
#pragma once

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

namespace medusa {

constexpr uint32_t kPrLane = 0;  // one lane per vertex, in-degree 1..64
constexpr uint32_t kPrWave = 1;  // one wave per vertex, in-degree > 64
constexpr uint32_t kPrZero = 2;  // no in-edges
constexpr uint32_t kPrLaneMax = 64;

// ---- binned message exchange (MEDUSA_PR_BINNED) ----------------------------------------
// A superstep in two launches.  Send: tile t of chunk c copies the chunk's contributions into
// LDS and writes one message per out-edge, four at a time, into its contiguous range of the
// message array.  Combine: workgroup b stages its bin's run table in LDS, reads its messages
// four at a time (a binary search over the table gives each group of four its run) together
// with their LDS slots, drops every message into its slot, and then sums each destination
// from LDS in the canonical order.  Destinations with more than kPrHubDeg in-edges do not fit
// a bin; they are summed by a gather over the CSC (as the pull variant does) in extra
// workgroups of the combine launch.
constexpr uint32_t kPrBinValues = 8192;       // LDS message slots of one combine workgroup
constexpr uint32_t kPrBinMaxVertices = 8192;  // destinations of one bin
constexpr uint32_t kPrBinMaxItems = 512;      // items of one bin (8 waves hold 64 records each)
constexpr uint32_t kPrBinRuns = 2048;         // runs (chunks sending to it) one bin may have
constexpr uint32_t kPrChunkSources = 8192;    // sources of one send chunk (their contributions fit LDS)
constexpr uint32_t kPrTileEdges = 16384;      // message entries written by one send workgroup (at least)
constexpr uint32_t kPrHubDeg = 4096;          // in-degree above which a destination is summed by gather
constexpr uint32_t kPrRunAlign = 4;           // runs start and end on multiples of this many entries
constexpr uint16_t kPrDummySlot = (uint16_t)kPrBinValues;  // LDS slot of padding entries
static_assert(kPrTileEdges % kPrRunAlign == 0, "tiles keep the alignment of their chunk");
static_assert(kPrBinValues < 65536 && kPrChunkSources <= 65536, "slots and source offsets are 16-bit");
static_assert(kPrHubDeg <= kPrBinValues, "a wave vertex must fit one bin");

// One combine item: a wave vertex (kPrWave; its k-th message at LDS lds+k), a group of up to
// 64 consecutive lane vertices (kPrLane; lane i's k-th message at lds + 64k + i), or up to 64
// consecutive vertices without in-edges (kPrZero).
struct PrItem {
    uint32_t v;      // first internal vertex
    uint32_t count;  // vertices (1 for kPrWave)
    uint32_t kind;
    uint32_t lds;    // LDS offset of the item's slots
    uint32_t part;   // packed graph of the item's vertices
    uint32_t deg;    // kPrWave: the in-degree; kPrLane: the largest in-degree of the group
    uint32_t pad[2];
};
// One combine workgroup: its items, its runs, and the LDS slots of its entries
// slot[slot_begin, slot_begin + runs' total length) in run order.
struct PrBin {
    uint32_t item_begin, item_end;
    uint32_t run_begin, run_end;
    uint32_t slot_begin;
    uint32_t entries;  // the bin's entries (its runs' total length, a multiple of kPrRunAlign)
    uint32_t pad[2];
};
// A run of entries from one chunk to one bin: `msg` = its first entry in the message array,
// `k` = its first entry's number among the bin's entries (both multiples of kPrRunAlign); the
// run's length is the next run's k minus this one's (the table ends each bin with a sentinel).
struct PrMsgRun {
    uint32_t msg, k;
};
// One send workgroup: entries [msg_begin, msg_end) of the chunk whose sources are
// [src_begin, src_begin + src_count); an entry's source is src_begin + srcloc[entry].
struct PrTile {
    uint32_t msg_begin, msg_end, src_begin, src_count;
};

struct PrBinsPlan {
    std::vector<PrItem> items;   // bin items, bin by bin, then one kPrWave item per hub
    std::vector<PrBin> bins;
    std::vector<uint32_t> bin_slots;  // LDS slots used by each bin
    std::vector<PrMsgRun> runs;  // per bin: its runs, then a sentinel {0, total length}
    std::vector<PrTile> tiles;
    std::vector<uint16_t> srcloc;  // message order: source offset in the chunk (0 for padding)
    std::vector<uint16_t> slot;    // bin order: LDS slot (kPrDummySlot for padding)
    std::vector<uint32_t> qmsg;    // bin order, per group of kPrRunAlign entries: its message-array
                                   // position / kPrRunAlign (a group never straddles two runs)
    uint32_t nhubs = 0;
    uint32_t max_runs = 0;  // most runs of any bin
    uint32_t tile_edges = kPrTileEdges;  // message entries per send tile (a chunk's last tile may hold fewer)
    uint64_t nmsg = 0;      // messages (in-edges of non-hub destinations)
    uint64_t nentry = 0;    // message-array entries, padding included
};

// Plans the binned superstep for a graph in the PageRank internal order: cptr[n+1] in-edge
// offsets, crow[m] in-edge sources (internal ids, canonical order per destination),
// part_begin[nparts+1] packed-graph ranges.  send_slots / combine_slots = send / combine
// workgroups that run at once on the device (0: unknown).  Returns "" or the reason the plan
// cannot be made.
inline std::string plan_pr_bins(uint32_t n, const uint32_t* cptr, const uint32_t* crow, const uint32_t* part_begin,
                                uint32_t nparts, PrBinsPlan& P, uint32_t send_slots = 0,
                                uint32_t combine_slots = 0) {
    P = PrBinsPlan{};
    const uint32_t kNone = 0xFFFFFFFFu;
    std::vector<uint32_t> bin_of(n, kNone), item_of(n, kNone);
    std::vector<PrItem> hubs;
    uint32_t item_begin = 0, used = 0, verts = 0, part = 0;
    int64_t group = -1;  // open lane / zero group
    uint32_t group_max = 0;
    uint32_t cap_values = kPrBinValues, cap_verts = kPrBinMaxVertices;  // of the bin being filled
    auto close_bin = [&] {
        if (P.items.size() > item_begin) {
            P.bins.push_back({item_begin, (uint32_t)P.items.size(), 0, 0, 0, 0, {0, 0}});
            P.bin_slots.push_back(used);
        }
        item_begin = (uint32_t)P.items.size();
        used = 0;
        verts = 0;
        group = -1;
    };
    auto full = [&](uint32_t foot) {
        return used + foot > cap_values || verts + 1 > cap_verts || P.items.size() - item_begin >= kPrBinMaxItems;
    };
    // Items and bins in vertex order; bins from vertex tail_from on hold a quarter of the slots
    // and vertices (an item larger than that gets a bin of its own, within kPrBinValues).
    auto plan_items = [&](uint32_t tail_from) {
        P.items.clear();
        P.bins.clear();
        P.bin_slots.clear();
        hubs.clear();
        std::fill(bin_of.begin(), bin_of.end(), kNone);
        std::fill(item_of.begin(), item_of.end(), kNone);
        item_begin = used = verts = part = 0;
        group = -1;
        group_max = 0;
        cap_values = kPrBinValues;
        cap_verts = kPrBinMaxVertices;
        for (uint32_t p = 0; p < n; ++p) {
            if (p == tail_from) {
                close_bin();
                cap_values = kPrBinValues / 4;
                cap_verts = kPrBinMaxVertices / 4;
            }
            while (part + 1 < nparts && p >= part_begin[part + 1]) {
                ++part;
                group = -1;
            }
            const uint32_t deg = cptr[p + 1] - cptr[p];
            if (deg > kPrHubDeg) {
                hubs.push_back({p, 1, kPrWave, 0, part, deg, {0, 0}});
                group = -1;
                continue;
            }
            const uint32_t kind = deg == 0 ? kPrZero : deg <= kPrLaneMax ? kPrLane : kPrWave;
            if (kind == kPrWave) {
                if (full(deg)) close_bin();
                item_of[p] = (uint32_t)P.items.size();
                P.items.push_back({p, 1, kPrWave, used, part, deg, {0, 0}});
                used += deg;
                group = -1;
            } else {
                const bool extend = group >= 0 && P.items[group].kind == kind && P.items[group].count < 64 &&
                                    P.items[group].v + P.items[group].count == p && deg <= group_max &&
                                    verts + 1 <= cap_verts;
                if (!extend) {
                    const uint32_t foot = kind == kPrLane ? 64u * deg : 0u;
                    if (full(foot)) close_bin();
                    group = (int64_t)P.items.size();
                    P.items.push_back({p, 0, kind, used, part, deg, {0, 0}});
                    group_max = deg;
                    used += foot;
                }
                ++P.items[group].count;
                item_of[p] = (uint32_t)group;
            }
            ++verts;
            bin_of[p] = (uint32_t)P.bins.size();
        }
        close_bin();
    };
    // The combine runs its bins in rounds of combine_slots workgroups.  When the last round is
    // short (at most a quarter of the slots: RMAT-20, 2231 bins for 1024), its bins would run
    // nearly alone at the end; the vertices of those bins are planned again into quarter-size
    // bins, which spread that work over the workgroups as they free up.
    plan_items(n);
    if (combine_slots && P.bins.size() > combine_slots) {
        const uint64_t r = P.bins.size() % combine_slots;
        if (r && r <= combine_slots / 4) plan_items(P.items[P.bins[P.bins.size() - r].item_begin].v);
    }
    const uint64_t B = P.bins.size();
    const uint64_t C = ((uint64_t)n + kPrChunkSources - 1) / kPrChunkSources;
    if (C * B > (1ull << 27)) return "binned PageRank: too many chunk x bin pairs for this graph";

    // messages per (chunk, bin)
    std::vector<uint32_t> cnt(C * B, 0u);
    for (uint32_t p = 0; p < n; ++p) {
        if (bin_of[p] == kNone) continue;
        const uint64_t b = bin_of[p];
        for (uint32_t q = cptr[p]; q < cptr[p + 1]; ++q) ++cnt[(crow[q] / kPrChunkSources) * B + b];
    }
    auto padded = [](uint64_t x) { return (x + kPrRunAlign - 1) / kPrRunAlign * kPrRunAlign; };
    // run starts in the message array (chunk-major) and in their bin's entry numbering
    std::vector<uint32_t> rstart(C * B), kstart(C * B), chunk_start(C + 1), slot_begin(B + 1);
    uint64_t acc = 0;
    for (uint64_t c = 0; c < C; ++c) {
        chunk_start[c] = (uint32_t)acc;
        for (uint64_t b = 0; b < B; ++b) {
            rstart[c * B + b] = (uint32_t)acc;
            acc += padded(cnt[c * B + b]);  // padding entries: source offset 0, dummy slot
        }
    }
    chunk_start[C] = (uint32_t)acc;
    if (acc >= (1ull << 32)) return "binned PageRank: message array too large for 32-bit offsets";
    P.nentry = acc;
    acc = 0;
    for (uint64_t b = 0; b < B; ++b) {
        slot_begin[b] = (uint32_t)acc;
        for (uint64_t c = 0; c < C; ++c) {
            kstart[c * B + b] = (uint32_t)(acc - slot_begin[b]);
            acc += padded(cnt[c * B + b]);
        }
    }
    slot_begin[B] = (uint32_t)acc;
    // fill, in canonical in-edge order inside every run
    P.srcloc.assign(P.nentry, 0);
    P.slot.assign(P.nentry, kPrDummySlot);
    std::vector<uint32_t> fill(C * B, 0u);
    for (uint32_t p = 0; p < n; ++p) {
        if (bin_of[p] == kNone) continue;
        const uint64_t b = bin_of[p];
        const PrItem& I = P.items[item_of[p]];
        for (uint32_t q = cptr[p]; q < cptr[p + 1]; ++q) {
            const uint32_t u = crow[q], k = q - cptr[p];
            const uint64_t cb = (u / kPrChunkSources) * B + b;
            const uint32_t i = fill[cb]++;
            P.srcloc[rstart[cb] + i] = (uint16_t)(u % kPrChunkSources);
            const uint32_t lds = I.kind == kPrWave ? I.lds + k : I.lds + 64u * k + (p - I.v);
            P.slot[slot_begin[b] + kstart[cb] + i] = (uint16_t)lds;
            ++P.nmsg;
        }
    }
    for (uint64_t b = 0; b < B; ++b) {
        PrBin& Bn = P.bins[b];
        Bn.run_begin = (uint32_t)P.runs.size();
        Bn.slot_begin = slot_begin[b];
        for (uint64_t c = 0; c < C; ++c)
            if (cnt[c * B + b]) P.runs.push_back({rstart[c * B + b], kstart[c * B + b]});
        Bn.run_end = (uint32_t)P.runs.size();
        if (Bn.run_end - Bn.run_begin > kPrBinRuns) return "binned PageRank: a bin receives messages from too many chunks";
        P.max_runs = std::max(P.max_runs, Bn.run_end - Bn.run_begin);
        Bn.entries = slot_begin[b + 1] - slot_begin[b];
        P.runs.push_back({0u, Bn.entries});  // sentinel: the bin's length
    }
    P.qmsg.assign(P.nentry / kPrRunAlign, 0u);
    for (uint64_t b = 0; b < B; ++b) {
        const PrBin& Bn = P.bins[b];
        for (uint32_t r = Bn.run_begin; r < Bn.run_end; ++r) {
            const PrMsgRun R = P.runs[r];
            const uint32_t e = r + 1 < Bn.run_end ? P.runs[r + 1].k : Bn.entries;
            for (uint32_t k = R.k; k < e; k += kPrRunAlign)
                P.qmsg[(Bn.slot_begin + k) / kPrRunAlign] = (R.msg + (k - R.k)) / kPrRunAlign;
        }
    }
    // Tiles of kPrTileEdges entries, or up to twice that when the tiles would overflow one round
    // of the device's resident send workgroups: a tile count a little above the slots leaves a
    // second round of a few tiles that runs on an otherwise idle device (RMAT-20: 1112 tiles for
    // 1024 slots).  With T entries per tile the chunks give at most nentry / T + (chunks that
    // send) tiles, so T = nentry / (slots - those chunks) fits one round.  A larger T is not
    // taken: tiles of many small chunks would then differ widely in length, the longest alone at
    // the end (a batch of 128 graphs, 789 chunks for 1023 slots, ran 17 % slower that way).
    // Tiles never get smaller: each recomputes its chunk's contributions.
    uint64_t sending = 0, fixed_tiles = 0;
    for (uint64_t c = 0; c < C; ++c) {
        sending += chunk_start[c + 1] > chunk_start[c];
        fixed_tiles += (chunk_start[c + 1] - chunk_start[c] + kPrTileEdges - 1) / kPrTileEdges;
    }
    if (send_slots > sending && fixed_tiles > send_slots) {
        const uint64_t even = (P.nentry + (send_slots - sending) - 1) / (send_slots - sending);
        const uint64_t T = (even + kPrRunAlign - 1) / kPrRunAlign * kPrRunAlign;
        if (T > kPrTileEdges && T <= 2ull * kPrTileEdges) P.tile_edges = (uint32_t)T;
    }
    for (uint64_t c = 0; c < C; ++c) {
        const uint32_t src = (uint32_t)(c * kPrChunkSources);
        const uint32_t cnt_src = std::min<uint32_t>(kPrChunkSources, n - src);
        for (uint64_t e = chunk_start[c]; e < chunk_start[c + 1]; e += P.tile_edges)
            P.tiles.push_back({(uint32_t)e, (uint32_t)std::min<uint64_t>(e + P.tile_edges, chunk_start[c + 1]), src,
                               cnt_src});
    }
    P.nhubs = (uint32_t)hubs.size();
    P.items.insert(P.items.end(), hubs.begin(), hubs.end());
    return "";
}

// Checks every index the binned kernels take from a plan against the buffers and LDS it
// addresses.  Returns "" or the first violation.
inline std::string check_pr_bins(uint32_t n, const uint32_t* cptr, const PrBinsPlan& P) {
    const uint64_t nb = P.bins.size();
    if (P.srcloc.size() != P.nentry || P.slot.size() != P.nentry || P.qmsg.size() * kPrRunAlign != P.nentry ||
        P.nentry % kPrRunAlign)
        return "bins: array sizes";
    uint64_t real = 0;
    for (uint64_t b = 0; b < nb; ++b) {
        const PrBin& B = P.bins[b];
        if (B.item_begin > B.item_end || B.item_end > P.items.size() - P.nhubs) return "bins: item range";
        if (B.item_end - B.item_begin > kPrBinMaxItems) return "bins: items per bin";
        if (B.run_begin > B.run_end || B.run_end >= P.runs.size() || B.run_end - B.run_begin > P.max_runs)
            return "bins: run range";
        if (P.bin_slots[b] > kPrBinValues) return "bins: LDS slots";
        const uint32_t K = P.runs[B.run_end].k;  // the sentinel
        if (K != B.entries) return "bins: entry count";
        if ((uint64_t)B.slot_begin + K > P.nentry || B.slot_begin % kPrRunAlign || K % kPrRunAlign)
            return "bins: slot range";
        if (B.run_begin < B.run_end && P.runs[B.run_begin].k != 0) return "bins: first run";
        for (uint32_t r = B.run_begin; r < B.run_end; ++r) {
            const PrMsgRun R = P.runs[r];
            const uint32_t e = r + 1 < B.run_end ? P.runs[r + 1].k : K;
            if (R.k >= e || R.msg % kPrRunAlign || R.k % kPrRunAlign || e % kPrRunAlign) return "bins: run shape";
            if ((uint64_t)R.msg + (e - R.k) > P.nentry) return "bins: run bounds";
            for (uint32_t k = R.k; k < e; k += kPrRunAlign)
                if ((uint64_t)P.qmsg[(B.slot_begin + k) / kPrRunAlign] * kPrRunAlign != (uint64_t)R.msg + (k - R.k))
                    return "bins: group position";
        }
        for (uint32_t i = 0; i < K; ++i) {
            const uint16_t t = P.slot[B.slot_begin + i];
            if (t == kPrDummySlot) continue;
            if (t >= P.bin_slots[b]) return "bins: slot beyond the bin";
            ++real;
        }
        uint32_t verts = 0;
        for (uint32_t i = B.item_begin; i < B.item_end; ++i) {
            const PrItem& I = P.items[i];
            if (I.count == 0 || I.count > 64 || (I.kind == kPrWave && I.count != 1)) return "bins: item count";
            if ((uint64_t)I.v + I.count > n) return "bins: item vertices";
            if (I.kind == kPrWave && I.deg != cptr[I.v + 1] - cptr[I.v]) return "bins: wave item degree";
            verts += I.count;
            for (uint32_t j = 0; j < I.count; ++j) {
                const uint32_t deg = cptr[I.v + j + 1] - cptr[I.v + j];
                if ((I.kind == kPrZero) != (deg == 0)) return "bins: zero item";
                if (I.kind == kPrLane && (deg == 0 || deg > kPrLaneMax || deg > I.deg)) return "bins: lane item degree";
            }
            const uint64_t top = I.kind == kPrWave ? (uint64_t)I.lds + I.deg
                                 : I.kind == kPrLane ? (uint64_t)I.lds + 64ull * I.deg
                                                     : 0;
            if (top > P.bin_slots[b]) return "bins: item beyond its LDS slots";
        }
        if (verts > kPrBinMaxVertices) return "bins: vertices per bin";
    }
    if (real != P.nmsg) return "bins: message count";
    if (P.tile_edges < kPrTileEdges || P.tile_edges % kPrRunAlign) return "bins: tile size";
    for (const PrTile& T : P.tiles) {
        if (T.msg_begin > T.msg_end || T.msg_end > P.nentry || T.msg_end - T.msg_begin > P.tile_edges)
            return "bins: tile entries";
        if (T.msg_begin % kPrRunAlign || T.msg_end % kPrRunAlign) return "bins: tile alignment";
        if (T.src_count == 0 || T.src_count > kPrChunkSources || (uint64_t)T.src_begin + T.src_count > n ||
            T.src_begin % 4)
            return "bins: tile sources";
        for (uint32_t e = T.msg_begin; e < T.msg_end; ++e)
            if (P.srcloc[e] >= T.src_count) return "bins: source beyond the chunk";
    }
    for (uint64_t h = P.items.size() - P.nhubs; h < P.items.size(); ++h)
        if (P.items[h].v >= n || P.items[h].kind != kPrWave) return "bins: hub item";
    return "";
}

}  // namespace medusa
