// Host build of the library's kNN search (avism_amd/csrc/knn_search.h): the
// k nearest (k-th distance and neighbour set, each neighbour once) and the
// fixed-radius ball, checked against brute force: random and lattice (exact
// ties) particle sets,
// periodic and not, k = 16, 21, 32, 47, 64, and boxes smaller than the search.
// Prints one line per case; exit status = number of failing cases.
#include <cmath>
#include <cstdio>
#include <random>
#include <algorithm>
#include <set>
#include <vector>

#include "../avism_amd/csrc/knn_search.h"

using namespace avism;

static int run(int n, int k, float L, int periodic, int lattice, unsigned seed, int nq, float rfac) {
  std::mt19937 rng(seed);
  std::uniform_real_distribution<float> U(-L / 2, L / 2);
  const int np = lattice ? n * n * n : n * n * n + 7;
  std::vector<float4> raw(np);
  for (int i = 0; i < np; ++i) {
    float4& p = raw[i];
    if (lattice) {
      const float d = L / n;
      const int a = i % n, b = (i / n) % n, c = i / (n * n);
      p.x = -L / 2 + (a + 0.5f) * d;
      p.y = -L / 2 + (b + 0.5f) * d;
      p.z = -L / 2 + (c + 0.5f) * d;
    } else {
      p.x = U(rng);
      p.y = U(rng);
      p.z = U(rng);
    }
    p.w = 1.0f;
  }
  Bins B;
  B.nb = n;
  B.lo = -L / 2;
  B.bsz = L / n;
  B.L = L;
  B.eps = L * 2e-6f + B.bsz * 1e-5f;
  B.periodic = periodic;
  std::vector<int> key(np), idx(np);
  for (int i = 0; i < np; ++i) {
    key[i] = B.bin_of(raw[i].x) + n * (B.bin_of(raw[i].y) + n * B.bin_of(raw[i].z));
    idx[i] = i;
  }
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return key[a] < key[b]; });
  std::vector<float4> pos(np);
  std::vector<uint32_t> orig(np);
  std::vector<int32_t> start(n * n * n + 1, 0);
  for (int i = 0; i < np; ++i) {
    pos[i] = raw[idx[i]];
    orig[i] = idx[i];
    start[key[idx[i]] + 1]++;
  }
  for (int b = 0; b < n * n * n; ++b) start[b + 1] += start[b];
  B.pos = pos.data();
  B.orig = orig.data();
  B.start = start.data();
  std::vector<uint32_t> keys(64), gmx(16);
  // first search radius: rfac x the k-th neighbour distance at the mean density
  const float r0 = rfac * std::cbrt(3.0f * k / (4.0f * 3.14159265f * np / (L * L * L)));
  int bad = 0;
  for (int qi = 0; qi < nq; ++qi) {
    float qx, qy, qz;
    if (qi % 2) {
      const float4 q = pos[(qi * 7919) % np];
      qx = q.x; qy = q.y; qz = q.z;
    } else {
      qx = U(rng); qy = U(rng); qz = U(rng);
    }
    Sel<1, 16> h;
    h.init(keys.data(), gmx.data(), k);
    bool scanned = false;
    std::multiset<uint32_t> got, ball;
    auto take = [&](int j, float, const float4&) { got.insert(orig[j]); };
    const float r2 = 0.6f * B.bsz * B.bsz * (1 + qi % 5);
    auto inball = [&](int j, float, const float4&) { ball.insert(orig[j]); };
    float dk2;
    if (periodic) {
      dk2 = knn_kth<1>(B, np, qx, qy, qz, r0, k, h, scanned);
      knn_each<1>(B, np, qx, qy, qz, k, dk2, scanned, h, PosLoad{B.pos}, take);
      ball_each<1>(B, np, qx, qy, qz, r2, PosLoad{B.pos}, inball);
    } else {
      dk2 = knn_kth<0>(B, np, qx, qy, qz, r0, k, h, scanned);
      knn_each<0>(B, np, qx, qy, qz, k, dk2, scanned, h, PosLoad{B.pos}, take);
      ball_each<0>(B, np, qx, qy, qz, r2, PosLoad{B.pos}, inball);
    }
    std::vector<std::pair<float, uint32_t>> all;
    std::multiset<uint32_t> want_ball;
    for (int j = 0; j < np; ++j) {
      const float d = dist2(pos[j].x, pos[j].y, pos[j].z, qx, qy, qz, L, periodic);
      all.push_back({d, orig[j]});
      if (d <= r2) want_ball.insert(orig[j]);
    }
    std::sort(all.begin(), all.end());
    std::multiset<uint32_t> want;
    for (int t = 0; t < k; ++t) want.insert(all[t].second);
    if (want != got || dk2 != all[k - 1].first || ball != want_ball) ++bad;
  }
  printf("n=%d k=%d periodic=%d lattice=%d r0=%.3g: %d/%d wrong\n", n, k, periodic, lattice, r0, bad, nq);
  return bad ? 1 : 0;
}

int main() {
  int fails = 0;
  // first radius well below, near and above the k-th distance (retry passes,
  // the usual single pass, pruning by the k-th distance alone)
  for (float rf : {0.3f, 1.3f, 3.0f})
    for (int per = 0; per < 2; ++per)
      for (int lat = 0; lat < 2; ++lat) {
        fails += run(20, 16, 40.0f, per, lat, 1, 400, rf);
        fails += run(16, 32, 32.0f, per, lat, 2, 400, rf);
        fails += run(12, 64, 30.0f, per, lat, 3, 200, rf);
        fails += run(5, 32, 10.0f, per, lat, 4, 200, rf);   // box smaller than the search
        fails += run(4, 16, 8.0f, per, lat, 5, 100, rf);
        fails += run(14, 21, 28.0f, per, lat, 6, 200, rf);  // k not a multiple of the group size
        fails += run(10, 47, 20.0f, per, lat, 7, 200, rf);
      }
  return fails;
}
