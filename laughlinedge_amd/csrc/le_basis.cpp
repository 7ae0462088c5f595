#include "le_basis.hpp"

#include <algorithm>

namespace le {

// Kelleher's accelAsc: every partition of n as an ascending list of parts.
std::vector<std::vector<int>> accel_asc(int n)
{
    std::vector<std::vector<int>> out;
    std::vector<int> a(n + 2, 0);
    int k = 1;
    int y = n - 1;
    while (k != 0) {
        int x = a[k - 1] + 1;
        k -= 1;
        while (2 * x <= y) {
            a[k] = x;
            y -= x;
            k += 1;
        }
        int l = k + 1;
        while (x <= y) {
            a[k] = x;
            a[l] = y;
            out.emplace_back(a.begin(), a.begin() + k + 2);
            x += 1;
            y -= 1;
        }
        a[k] = x + y;
        y = x + y - 1;
        out.emplace_back(a.begin(), a.begin() + k + 1);
    }
    return out;
}

static std::vector<PartPair> compress(const std::vector<int>& parts)
{
    std::vector<PartPair> v;
    int lo = *std::min_element(parts.begin(), parts.end());
    int hi = *std::max_element(parts.begin(), parts.end());
    for (int x = lo; x <= hi; ++x) {
        int c = (int)std::count(parts.begin(), parts.end(), x);
        if (c) v.push_back({x, c});
    }
    // stable: only strictly larger multiplicities move right
    for (size_t i = 0; i < v.size(); ++i)
        for (size_t j = 0; j + i + 1 < v.size(); ++j)
            if (v[j].repetitions > v[j + 1].repetitions) std::swap(v[j], v[j + 1]);
    return v;
}

void Basis::build(int n_particles, const std::vector<int>& dls)
{
    N = n_particles;
    sectors = dls;
    states.clear();
    sector_offset.assign(1, 0);
    K = 0;
    longest = 0;
    for (int dl : dls) {
        for (const auto& p : accel_asc(dl)) {
            if ((int)p.size() > N) continue;
            BasisState st{dl, compress(p)};
            for (const auto& pp : st.pairs) K = std::max(K, pp.number);
            longest = std::max(longest, (int)st.pairs.size());
            states.push_back(st);
        }
        sector_offset.push_back((int)states.size());
    }
    D = (int)states.size();
}

long Basis::block_offset(int s) const
{
    long off = 0;
    for (int q = 0; q < s; ++q) off += (long)dim(q) * dim(q);
    return off;
}

long Basis::block_elems() const { return block_offset(nsec()); }

std::vector<uint32_t> Basis::factor_table() const
{
    std::vector<uint32_t> t((size_t)D * kMaxFactors, 0u);
    for (int s = 0; s < D; ++s) {
        int f = 0;
        for (const auto& pp : states[s].pairs) {
            if (pp.number == 0) continue;   // empty partition: constant 1
            t[(size_t)s * kMaxFactors + f++] = (uint32_t)pp.number | ((uint32_t)pp.repetitions << 8);
        }
    }
    return t;
}

}  // namespace le
