// ygpu_gtable.h -- launch table of member groups of different sizes (plain
// C++, no HIP: the host builds it, the group kernel reads it).
//
// Member k has group size mgroup[k] (a power of two, at most per_xcd).  A
// workgroup b of the launch runs on XCD b % 8 at slot b / 8; an XCD holds
// per_xcd slots (workgroups resident at once), a multiple of 16.  The table:
//   gtab[3 b + {0, 1, 2}] = group, rank, group size of workgroup b (group -1:
//                           no part, the workgroup returns at once);
//   group g takes the members gmem[gmoff[g] .. gmoff[g+1]) in turn;
//   nwg = 8 x (the most slots any XCD uses).
// Rules: one group per member when all groups fit the 8 x per_xcd slots at
// once, else fewer groups of the size that holds the most workgroups (its
// members then share them in turn, member q of the size going to group
// q mod groups); groups are placed largest first, each on the XCD with the
// most free slots (the first of equals), on consecutive slots.  Sizes are
// powers of two and per_xcd a multiple of the largest, so every XCD's free
// slots stay a multiple of the next size and the placement fails only when
// the slots are all taken; groups of one size go round the XCDs in turn
// (uniform groups: group g on XCD g % 8 at slots G (g / 8) ..., the kernel's
// own placement without a table).
#pragma once
#include <algorithm>
#include <functional>
#include <map>
#include <string>
#include <utility>
#include <vector>

namespace yg {

inline bool group_table(const std::vector<int>& mgroup, int per_xcd, std::vector<int>& gtab,
                        std::vector<int>& gmoff, std::vector<int>& gmem, int& nwg,
                        std::string& err) {
  const int nmem = (int)mgroup.size();
  const long long cap = 8LL * per_xcd;
  std::map<int, std::vector<int>, std::greater<int>> cls;   // size -> members
  for (int k = 0; k < nmem; ++k) {
    const int g = mgroup[k];
    if (g < 1 || g > per_xcd || (g & (g - 1))) { err = "member groups: bad group size"; return false; }
    cls[g].push_back(k);
  }
  std::map<int, int> kc;   // groups per size
  long long used = 0;
  for (auto& c : cls) {
    kc[c.first] = (int)c.second.size();
    used += (long long)c.first * kc[c.first];
  }
  while (used > cap) {
    int best = -1;
    long long bw = 0;
    for (auto& c : kc)
      if (c.second > 1 && (long long)c.first * c.second > bw) {
        bw = (long long)c.first * c.second;
        best = c.first;
      }
    if (best < 0) { err = "member groups: the groups of the batch exceed the CUs"; return false; }
    --kc[best];
    used -= best;
  }
  std::vector<std::pair<int, std::vector<int>>> grp;   // largest first
  for (auto& c : cls) {
    const int n = kc[c.first];
    for (int j = 0; j < n; ++j) {
      std::vector<int> ms;
      for (size_t q = j; q < c.second.size(); q += n) ms.push_back(c.second[q]);
      grp.push_back({c.first, ms});
    }
  }
  std::vector<int> fill(8, 0);
  std::vector<std::pair<int, int>> place(grp.size());   // (XCD, first slot)
  for (size_t g = 0; g < grp.size(); ++g) {
    int x = 0;   // the XCD with the most free slots, the first of equals
    for (int y = 1; y < 8; ++y)
      if (fill[y] < fill[x]) x = y;
    if (fill[x] + grp[g].first > per_xcd) { err = "member groups: no XCD holds a group"; return false; }
    place[g] = {x, fill[x]};
    fill[x] += grp[g].first;
  }
  nwg = 8 * *std::max_element(fill.begin(), fill.end());
  gtab.assign(3 * (size_t)nwg, 0);
  for (int b = 0; b < nwg; ++b) gtab[3 * b] = -1;
  gmoff.assign(grp.size() + 1, 0);
  gmem.clear();
  for (size_t g = 0; g < grp.size(); ++g) {
    for (int r = 0; r < grp[g].first; ++r) {
      const int b = 8 * (place[g].second + r) + place[g].first;
      gtab[3 * b] = (int)g;
      gtab[3 * b + 1] = r;
      gtab[3 * b + 2] = grp[g].first;
    }
    gmem.insert(gmem.end(), grp[g].second.begin(), grp[g].second.end());
    gmoff[g + 1] = (int)gmem.size();
  }
  return true;
}

}  // namespace yg
