// Edge basis: partitions of each requested angular-momentum sector dL with at most N
// parts, compressed to (part, multiplicity) pairs.
//
// Order and compression follow modules/partitions/integer_partitions.cu:
//   accelAsc enumeration ............................ :10-47
//   keep partitions with <= N parts ................. :104-117
//   (part, multiplicity), increasing part ........... :118-131
//   stable sort by increasing multiplicity .......... :63-77
//   flatten over sectors in the order requested ..... :199-226
// dL = 0 gives the single pair (0, 1); its polynomial is the constant 1 here.
#pragma once
#include <cstdint>
#include <vector>

namespace le {

struct PartPair {
    int number;
    int repetitions;
};

struct BasisState {
    int dL;
    std::vector<PartPair> pairs;
};

constexpr int kMaxFactors = 8;   // compressed length bound handled by the GPU tables

struct Basis {
    int N = 0;
    std::vector<int> sectors;          // requested dL values
    std::vector<BasisState> states;    // flattened
    std::vector<int> sector_offset;    // nsec + 1
    int D = 0;
    int K = 0;                         // largest part
    int longest = 0;                   // longest compressed length

    void build(int n_particles, const std::vector<int>& dls);
    int nsec() const { return (int)sectors.size(); }
    int dim(int s) const { return sector_offset[s + 1] - sector_offset[s]; }
    long block_elems() const;          // sum_s d_s^2
    long block_offset(int s) const;    // sum_{q<s} d_q^2
    // factor table: per state kMaxFactors packed words (k | r << 8), 0 terminates
    std::vector<uint32_t> factor_table() const;
};

std::vector<std::vector<int>> accel_asc(int n);

}  // namespace le
