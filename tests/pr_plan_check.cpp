// Host-side checker for the binned PageRank planner (medusa_amd/csrc/pr_bins.h), used by
// tests/test_pr_plan.py: plans a graph given as raw CSC files with and without the device's
// resident-workgroup counts and prints the plan's shape as one JSON line.
//
//   pr_plan_check <cptr.u32> <crow.u32> <part_begin.u32> <send_slots> <combine_slots>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "pr_bins.h"

using namespace medusa;

static std::vector<uint32_t> read_u32(const char* path) {
    std::vector<uint32_t> v;
    FILE* f = std::fopen(path, "rb");
    if (!f) return v;
    uint32_t x;
    while (std::fread(&x, 4, 1, f) == 1) v.push_back(x);
    std::fclose(f);
    return v;
}

int main(int argc, char** argv) {
    if (argc != 6) return 2;
    const std::vector<uint32_t> cptr = read_u32(argv[1]), crow = read_u32(argv[2]), pb = read_u32(argv[3]);
    const uint32_t send_slots = (uint32_t)std::strtoul(argv[4], nullptr, 10);
    const uint32_t combine_slots = (uint32_t)std::strtoul(argv[5], nullptr, 10);
    if (cptr.empty() || pb.size() < 2) return 2;
    const uint32_t n = (uint32_t)cptr.size() - 1;
    PrBinsPlan P0, P1;
    const std::string e0 = plan_pr_bins(n, cptr.data(), crow.data(), pb.data(), (uint32_t)pb.size() - 1, P0);
    const std::string e1 = plan_pr_bins(n, cptr.data(), crow.data(), pb.data(), (uint32_t)pb.size() - 1, P1,
                                        send_slots, combine_slots);
    if (!e0.empty() || !e1.empty()) {
        std::printf("{\"err\": \"%s%s\"}\n", e0.c_str(), e1.c_str());
        return 0;
    }
    const std::string c0 = check_pr_bins(n, cptr.data(), P0), c1 = check_pr_bins(n, cptr.data(), P1);
    uint32_t max_tile = 0;
    for (const PrTile& T : P1.tiles) max_tile = std::max(max_tile, T.msg_end - T.msg_begin);
    // bins of plan 1 that lie past plan 0's last whole round of combine_slots bins
    const size_t whole = combine_slots ? P0.bins.size() / combine_slots * combine_slots : P0.bins.size();
    uint32_t tail_bins = 0, tail_over_quarter = 0, tail_max_slots = 0, tail_over_quarter_shared = 0;
    for (size_t b = whole; b < P1.bins.size(); ++b) {
        ++tail_bins;
        tail_max_slots = std::max(tail_max_slots, P1.bin_slots[b]);
        if (P1.bin_slots[b] > kPrBinValues / 4) {
            ++tail_over_quarter;
            if (P1.bins[b].item_end - P1.bins[b].item_begin > 1) ++tail_over_quarter_shared;
        }
    }
    std::printf("{\"err\": \"\", \"check0\": \"%s\", \"check1\": \"%s\", \"bins0\": %zu, \"bins1\": %zu, "
                "\"tiles0\": %zu, \"tiles1\": %zu, \"tile_edges1\": %u, \"max_tile1\": %u, \"nmsg0\": %llu, "
                "\"nmsg1\": %llu, \"tail_bins\": %u, \"tail_over_quarter\": %u, \"tail_over_quarter_shared\": %u, "
                "\"tail_max_slots\": %u}\n",
                c0.c_str(), c1.c_str(), P0.bins.size(), P1.bins.size(), P0.tiles.size(), P1.tiles.size(),
                P1.tile_edges, max_tile, (unsigned long long)P0.nmsg, (unsigned long long)P1.nmsg, tail_bins,
                tail_over_quarter, tail_over_quarter_shared, tail_max_slots);
    return 0;
}
