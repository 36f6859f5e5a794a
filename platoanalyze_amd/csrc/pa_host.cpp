// Host-side setup: box mesh (Omega_h::build_box numbering), mesh sets, block graph,
// slot codes and LDS chunks.  See pa_host.hpp.
#include <algorithm>
#include <cmath>
#include <functional>
#include <numeric>
#include <thread>
#include <unordered_map>

#include "pa_host.hpp"

namespace pa {

void parallel_for(int64_t n, const std::function<void(int64_t, int64_t)>& body) {
  if (n <= 0) return;
  unsigned hw = std::thread::hardware_concurrency();
  int nt = (int)std::max(1u, std::min(hw ? hw : 1u, 16u));
  if (n < 4096 || nt == 1) {
    body(0, n);
    return;
  }
  std::vector<std::thread> th;
  int64_t step = (n + nt - 1) / nt;
  for (int t = 0; t < nt; ++t) {
    int64_t b = t * step, e = std::min(n, b + step);
    if (b >= e) break;
    th.emplace_back([&body, b, e] { body(b, e); });
  }
  for (auto& t : th) t.join();
}

// ---------------------------------------------------------------------------
// Hilbert key: John Skilling, "Programming the Hilbert curve", AIP Conf. Proc.
// 707 (2004) 381.  Axes -> transpose form, then the transposed bits are
// interleaved into a 3-word key (most significant word first).  Coordinates are
// scaled by the largest bounding-box extent onto a 52-bit grid; the upper end of
// the box is clamped to the last grid cell.  This reproduces the vertex numbering
// that the reference's box meshes carry (pinned by unit_tests/DerivativeTests.cpp
// residual/grad_z goldens on the 2x2x2 box).
static inline void axes_to_transpose(uint64_t* X, int b) {
  const int n = 3;
  uint64_t M = uint64_t(1) << (b - 1), P, Q, t;
  for (Q = M; Q > 1; Q >>= 1) {
    P = Q - 1;
    for (int i = 0; i < n; ++i) {
      if (X[i] & Q) {
        X[0] ^= P;
      } else {
        t = (X[0] ^ X[i]) & P;
        X[0] ^= t;
        X[i] ^= t;
      }
    }
  }
  for (int i = 1; i < n; ++i) X[i] ^= X[i - 1];
  t = 0;
  for (Q = M; Q > 1; Q >>= 1)
    if (X[n - 1] & Q) t ^= Q - 1;
  for (int i = 0; i < n; ++i) X[i] ^= t;
}

struct Key3 {
  uint64_t w[3];
  bool operator<(const Key3& o) const {
    if (w[0] != o.w[0]) return w[0] < o.w[0];
    if (w[1] != o.w[1]) return w[1] < o.w[1];
    return w[2] < o.w[2];
  }
};

static inline Key3 untranspose(const uint64_t* X) {
  Key3 k{{0, 0, 0}};
  for (int i = 0; i < 64 * 3; ++i) {
    int bit_in = 63 - i / 3, in_word = i % 3;
    int bit_out = 63 - (i % 64), out_word = i / 64;
    if (X[in_word] & (uint64_t(1) << bit_in)) k.w[out_word] |= uint64_t(1) << bit_out;
  }
  return k;
}

std::vector<int32_t> hilbert_order(const double* c, int n) {
  double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int i = 0; i < n; ++i)
    for (int d = 0; d < 3; ++d) {
      mn[d] = std::min(mn[d], c[3 * i + d]);
      mx[d] = std::max(mx[d], c[3 * i + d]);
    }
  double maxl = 0;
  for (int d = 0; d < 3; ++d) maxl = std::max(maxl, mx[d] - mn[d]);
  const int nbits = 52;
  const uint64_t top = (uint64_t(1) << nbits) - 1;
  std::vector<Key3> keys(n);
  parallel_for(n, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) {
      uint64_t X[3];
      for (int d = 0; d < 3; ++d) {
        double f = maxl > 0 ? (c[3 * i + d] - mn[d]) / maxl : 0.0;
        double s = std::ldexp(f, nbits);
        uint64_t v = s <= 0 ? 0 : (uint64_t)s;
        X[d] = std::min(v, top);
      }
      axes_to_transpose(X, nbits);
      keys[i] = untranspose(X);
    }
  });
  std::vector<int32_t> perm(n);
  std::iota(perm.begin(), perm.end(), 0);
  std::stable_sort(perm.begin(), perm.end(), [&](int32_t a, int32_t b) { return keys[a] < keys[b]; });
  return perm;
}

// Six tets per hex around the v0-v6 diagonal, local vertex order as the reference
// box meshes carry it (unit_tests/OmegaHUtilitiesTests.cpp LocalElementCoords_3D).
// Hex local vertices: 0 (0,0,0) 1 (1,0,0) 2 (1,1,0) 3 (0,1,0) 4 (0,0,1) 5 (1,0,1)
// 6 (1,1,1) 7 (0,1,1).
static const int kHexTets[6][4] = {{0, 2, 3, 6}, {0, 3, 7, 6}, {0, 7, 4, 6},
                                   {0, 5, 6, 4}, {1, 5, 6, 0}, {1, 6, 2, 0}};

BoxMesh build_box(int nx, int ny, int nz, double lx, double ly, double lz) {
  BoxMesh m;
  const int nvx = nx + 1, nvy = ny + 1, nvz = nz + 1;
  const int64_t nvxy = (int64_t)nvx * nvy;
  const int64_t nv = nvxy * nvz;
  const int64_t nh = (int64_t)nx * ny * nz;
  const double dx = lx / nx, dy = ly / ny, dz = lz / nz;
  std::vector<double> c0(3 * nv);
  parallel_for(nv, [&](int64_t b, int64_t e) {
    for (int64_t v = b; v < e; ++v) {
      int64_t k = v / nvxy, r = v % nvxy, j = r / nvx, i = r % nvx;
      c0[3 * v + 0] = i * dx;
      c0[3 * v + 1] = j * dy;
      c0[3 * v + 2] = k * dz;
    }
  });
  std::vector<int32_t> perm = hilbert_order(c0.data(), (int)nv);  // new -> old
  std::vector<int32_t> o2n(nv);
  for (int64_t i = 0; i < nv; ++i) o2n[perm[i]] = (int32_t)i;
  m.num_nodes = (int)nv;
  m.coords.resize(3 * nv);
  parallel_for(nv, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i)
      for (int d = 0; d < 3; ++d) m.coords[3 * i + d] = c0[3 * (int64_t)perm[i] + d];
  });
  const int64_t ne = nh * 6;
  std::vector<int32_t> conn0(4 * ne);
  parallel_for(nh, [&](int64_t b, int64_t e) {
    for (int64_t q = b; q < e; ++q) {
      int64_t k = q / ((int64_t)nx * ny), r = q % ((int64_t)nx * ny), j = r / nx, i = r % nx;
      int64_t vv = k * nvxy + j * nvx + i;
      int64_t hv[8] = {vv, vv + 1, vv + nvx + 1, vv + nvx, vv + nvxy, vv + nvxy + 1,
                       vv + nvxy + nvx + 1, vv + nvxy + nvx};
      for (int t = 0; t < 6; ++t)
        for (int a = 0; a < 4; ++a) conn0[4 * (6 * q + t) + a] = o2n[hv[kHexTets[t][a]]];
    }
  });
  // Elements ordered by smallest vertex (stable): groups each element with the
  // chunk of rows that owns its first vertex, which keeps list reads contiguous.
  std::vector<int32_t> minv(ne);
  for (int64_t e = 0; e < ne; ++e) {
    int32_t v = conn0[4 * e];
    for (int a = 1; a < 4; ++a) v = std::min(v, conn0[4 * e + a]);
    minv[e] = v;
  }
  // counting sort by min vertex (stable)
  std::vector<int64_t> cnt(nv + 1, 0);
  for (int64_t e = 0; e < ne; ++e) cnt[minv[e] + 1]++;
  for (int64_t i = 0; i < nv; ++i) cnt[i + 1] += cnt[i];
  m.conn.resize(4 * ne);
  for (int64_t e = 0; e < ne; ++e) {
    int64_t pos = cnt[minv[e]]++;
    for (int a = 0; a < 4; ++a) m.conn[4 * pos + a] = conn0[4 * e + a];
  }
  m.num_elems = (int)ne;
  return m;
}

std::vector<int32_t> nodes_on_plane(const double* c, int nn, int axis, double value, double tol) {
  std::vector<int32_t> out;
  for (int i = 0; i < nn; ++i)
    if (std::fabs(c[3 * i + axis] - value) <= tol) out.push_back(i);
  return out;
}

// Tet faces listed by the vertex they omit; a face lying on a box plane belongs to
// exactly one tet, so no deduplication is needed for boundary planes.
void faces_on_plane(const int32_t* conn, int ne, const double* c, int axis, double value, double tol,
                    std::vector<int32_t>& fv, std::vector<int32_t>& fe) {
  static const int kFace[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
  fv.clear();
  fe.clear();
  for (int e = 0; e < ne; ++e) {
    bool on[4];
    int non = 0;
    for (int a = 0; a < 4; ++a) {
      on[a] = std::fabs(c[3 * conn[4 * e + a] + axis] - value) <= tol;
      non += on[a];
    }
    if (non < 3) continue;
    for (int f = 0; f < 4; ++f) {
      if (on[kFace[f][0]] && on[kFace[f][1]] && on[kFace[f][2]]) {
        for (int k = 0; k < 3; ++k) fv.push_back(conn[4 * e + kFace[f][k]]);
        fe.push_back(e);
      }
    }
  }
}

NodeElems node_elems(const int32_t* conn, int ne, int nn) {
  NodeElems r;
  r.offs.assign(nn + 1, 0);
  for (int64_t k = 0; k < 4LL * ne; ++k) r.offs[conn[k] + 1]++;
  for (int i = 0; i < nn; ++i) r.offs[i + 1] += r.offs[i];
  r.ent.resize(4LL * ne);
  std::vector<int32_t> pos(r.offs.begin(), r.offs.end() - 1);
  for (int64_t e = 0; e < ne; ++e)
    for (int a = 0; a < 4; ++a) r.ent[pos[conn[4 * e + a]]++] = (int32_t)(e * 4 + a);
  return r;
}

Graph build_graph(const int32_t* conn, int ne, int nn) {
  NodeElems adj = node_elems(conn, ne, nn);
  Graph g;
  g.num_rows = nn;
  std::vector<std::vector<int32_t>> rows(nn);
  parallel_for(nn, [&](int64_t b, int64_t e) {
    std::vector<int32_t> tmp;
    for (int64_t r = b; r < e; ++r) {
      tmp.clear();
      tmp.push_back((int32_t)r);
      for (int k = adj.offs[r]; k < adj.offs[r + 1]; ++k) {
        int el = adj.ent[k] >> 2;
        for (int a = 0; a < 4; ++a) tmp.push_back(conn[4 * el + a]);
      }
      std::sort(tmp.begin(), tmp.end());
      tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
      rows[r] = tmp;
    }
  });
  g.rowmap.resize(nn + 1);
  g.rowmap[0] = 0;
  for (int r = 0; r < nn; ++r) {
    g.rowmap[r + 1] = g.rowmap[r] + (int32_t)rows[r].size();
    g.max_row_len = std::max(g.max_row_len, (int)rows[r].size());
  }
  g.colidx.resize(g.rowmap[nn]);
  parallel_for(nn, [&](int64_t b, int64_t e) {
    for (int64_t r = b; r < e; ++r) std::copy(rows[r].begin(), rows[r].end(), g.colidx.begin() + g.rowmap[r]);
  });
  return g;
}

std::string check_graph(Graph& g, const int32_t* conn, int ne, int nn) {
  if ((int)g.rowmap.size() != nn + 1) return "graph rowmap must have num_nodes+1 entries";
  if (g.rowmap[0] != 0) return "graph rowmap[0] must be 0";
  for (int r = 0; r < nn; ++r)
    if (g.rowmap[r + 1] < g.rowmap[r]) return "graph rowmap must be non-decreasing";
  if ((int64_t)g.colidx.size() != g.rowmap[nn]) return "graph colidx size must equal rowmap[num_nodes]";
  g.num_rows = nn;
  g.max_row_len = 0;
  for (int r = 0; r < nn; ++r) {
    g.max_row_len = std::max(g.max_row_len, g.rowmap[r + 1] - g.rowmap[r]);
    bool diag = false;
    for (int k = g.rowmap[r]; k < g.rowmap[r + 1]; ++k) {
      if (g.colidx[k] < 0 || g.colidx[k] >= nn) return "graph column index out of range";
      diag |= g.colidx[k] == r;
    }
    if (!diag) return "graph row " + std::to_string(r) + " has no diagonal entry";
  }
  (void)conn;
  (void)ne;
  return "";
}

std::vector<uint32_t> slot_codes(const Graph& g, const int32_t* conn, int ne, int bits, std::string& err) {
  const int words = bits == 4 ? 2 : 4;  // 16 pairs * bits / 32
  std::vector<uint32_t> codes((size_t)ne * words, 0u);
  std::vector<int> bad(1, -1);
  parallel_for(ne, [&](int64_t b, int64_t e) {
    for (int64_t el = b; el < e; ++el) {
      for (int a = 0; a < 4; ++a) {
        int ra = conn[4 * el + a];
        int s0 = g.rowmap[ra], s1 = g.rowmap[ra + 1];
        for (int bb = 0; bb < 4; ++bb) {
          int cb = conn[4 * el + bb];
          int pos = -1;
          for (int k = s0; k < s1; ++k)
            if (g.colidx[k] == cb) {
              pos = k - s0;
              break;
            }
          if (pos < 0 || pos >= (1 << bits)) {
            bad[0] = (int)el;
            continue;
          }
          int pair = 4 * a + bb;
          int bitpos = pair * bits;
          codes[(size_t)el * words + bitpos / 32] |= (uint32_t)pos << (bitpos % 32);
        }
      }
    }
  });
  if (bad[0] >= 0)
    err = "element " + std::to_string(bad[0]) + " has a vertex pair missing from the graph (or row too long)";
  return codes;
}

Chunks build_chunks(const Graph& g, const NodeElems& adj, const int32_t* conn, const std::vector<uint32_t>& codes,
                    int wpc, int max_blocks, int max_rows, int max_nodes) {
  Chunks ch;
  const int nn = g.num_rows;
  // greedy partition: add rows while blocks, rows and node-table size fit
  {
    std::vector<int32_t> mark(nn, -1);
    ch.row_begin.push_back(0);
    int r = 0, cid = 0;
    while (r < nn) {
      const int r0 = r, b0 = g.rowmap[r0];
      int nodes = 0;
      while (r < nn && (r - r0) < max_rows && (g.rowmap[r + 1] - b0) <= max_blocks) {
        int add = 0;
        for (int k = g.rowmap[r]; k < g.rowmap[r + 1]; ++k)
          if (mark[g.colidx[k]] != cid) ++add;
        if (nodes + add > max_nodes && r > r0) break;
        for (int k = g.rowmap[r]; k < g.rowmap[r + 1]; ++k)
          if (mark[g.colidx[k]] != cid) mark[g.colidx[k]] = cid;
        nodes += add;
        ++r;
      }
      if (r == r0) ++r;  // one row over budget still forms a chunk; the caller rejects it
      ch.row_begin.push_back(r);
      ++cid;
    }
  }
  ch.count = (int)ch.row_begin.size() - 1;
  std::vector<std::vector<uint32_t>> lists(ch.count);
  std::vector<std::vector<int32_t>> tables(ch.count);
  std::vector<std::vector<uint16_t>> lns(ch.count);
  std::vector<int> mrows(ch.count), mblocks(ch.count), mnodes(ch.count);
  parallel_for(ch.count, [&](int64_t b, int64_t e) {
    std::vector<uint64_t> tmp;
    std::vector<int32_t> local;  // global -> local, lazily sized
    std::unordered_map<int32_t, int32_t> lmap;
    for (int64_t c = b; c < e; ++c) {
      const int r0 = ch.row_begin[c], r1 = ch.row_begin[c + 1];
      // node table: rows first, then the other coupled nodes ascending
      auto& T = tables[c];
      T.clear();
      for (int rr = r0; rr < r1; ++rr) T.push_back(rr);
      std::vector<int32_t> halo;
      for (int rr = r0; rr < r1; ++rr)
        for (int k = g.rowmap[rr]; k < g.rowmap[rr + 1]; ++k)
          if (g.colidx[k] < r0 || g.colidx[k] >= r1) halo.push_back(g.colidx[k]);
      std::sort(halo.begin(), halo.end());
      halo.erase(std::unique(halo.begin(), halo.end()), halo.end());
      T.insert(T.end(), halo.begin(), halo.end());
      lmap.clear();
      for (size_t i = 0; i < T.size(); ++i) lmap[T[i]] = (int32_t)i;
      mrows[c] = r1 - r0;
      mblocks[c] = g.rowmap[r1] - g.rowmap[r0];
      mnodes[c] = (int)T.size();
      // entries
      tmp.clear();
      for (int rr = r0; rr < r1; ++rr)
        for (int k = adj.offs[rr]; k < adj.offs[rr + 1]; ++k) {
          uint64_t el = (uint64_t)(adj.ent[k] >> 2);
          int a = adj.ent[k] & 3;
          tmp.push_back((el << 4) | (1u << a));
        }
      std::sort(tmp.begin(), tmp.end());
      auto& L = lists[c];
      L.clear();
      for (size_t i = 0; i < tmp.size();) {
        uint64_t el = tmp[i] >> 4;
        uint32_t mask = 0;
        while (i < tmp.size() && (tmp[i] >> 4) == el) mask |= (uint32_t)(tmp[i++] & 15);
        L.push_back((uint32_t)el | (mask << 28));
      }
      // Group by owned-vertex mask (most-owned first, element order within a group):
      // the lanes of a wavefront then skip the same element edges (those with no owned
      // end), so a skipped edge costs nothing instead of running masked.
      std::stable_sort(L.begin(), L.end(), [](uint32_t x, uint32_t y) {
        const int px = __builtin_popcount(x >> 28), py = __builtin_popcount(y >> 28);
        return px != py ? px > py : (x >> 28) < (y >> 28);
      });
      // Within each group, interleave with stride ceil(n/64): the 64 lanes of a
      // wavefront then take elements spread over the chunk instead of neighbours
      // sharing vertices, so their LDS atomic adds rarely hit the same address.
      std::vector<uint32_t> tmpL(L.size());
      size_t g0 = 0;
      while (g0 < L.size()) {
        size_t g1 = g0;
        const uint32_t gm = L[g0] >> 28;
        while (g1 < L.size() && (L[g1] >> 28) == gm) ++g1;
        const size_t n = g1 - g0, q = (n + 63) / 64;
        size_t k = 0;
        for (size_t col = 0; col < q; ++col)
          for (size_t row = col; row < n; row += q) tmpL[g0 + k++] = L[g0 + row];
        g0 = g1;
      }
      L.swap(tmpL);
      auto& LN = lns[c];
      LN.resize(4 * L.size());
      for (size_t i = 0; i < L.size(); ++i) {
        const int el = (int)(L[i] & 0x0FFFFFFFu);
        for (int a = 0; a < 4; ++a) LN[4 * i + a] = (uint16_t)lmap.at(conn[4 * el + a]);
      }
    }
  });
  ch.list_offs.assign(ch.count + 1, 0);
  ch.node_offs.assign(ch.count + 1, 0);
  for (int c = 0; c < ch.count; ++c) {
    ch.list_offs[c + 1] = ch.list_offs[c] + (int32_t)lists[c].size();
    ch.node_offs[c + 1] = ch.node_offs[c] + (int32_t)tables[c].size();
    ch.max_rows = std::max(ch.max_rows, mrows[c]);
    ch.max_blocks = std::max(ch.max_blocks, mblocks[c]);
    ch.max_nodes = std::max(ch.max_nodes, mnodes[c]);
  }
  ch.list.resize(ch.list_offs[ch.count]);
  ch.nodes.resize(ch.node_offs[ch.count]);
  ch.ent_ln.resize(4 * (size_t)ch.list_offs[ch.count]);
  ch.ent_code.resize((size_t)wpc * ch.list_offs[ch.count]);
  parallel_for(ch.count, [&](int64_t b, int64_t e) {
    for (int64_t c = b; c < e; ++c) {
      std::copy(lists[c].begin(), lists[c].end(), ch.list.begin() + ch.list_offs[c]);
      std::copy(tables[c].begin(), tables[c].end(), ch.nodes.begin() + ch.node_offs[c]);
      std::copy(lns[c].begin(), lns[c].end(), ch.ent_ln.begin() + 4 * (size_t)ch.list_offs[c]);
      for (size_t i = 0; i < lists[c].size(); ++i) {
        const size_t el = lists[c][i] & 0x0FFFFFFFu;
        for (int wd = 0; wd < wpc; ++wd)
          ch.ent_code[(size_t)wpc * (ch.list_offs[c] + i) + wd] = codes[(size_t)wpc * el + wd];
      }
    }
  });
  ch.tslot.assign(g.colidx.size(), kNoSlot);
  parallel_for(ch.count, [&](int64_t b, int64_t e) {
    for (int64_t c = b; c < e; ++c) {
      const int r0 = ch.row_begin[c], r1 = ch.row_begin[c + 1], b0 = g.rowmap[r0];
      for (int r = r0; r < r1; ++r)
        for (int k = g.rowmap[r]; k < g.rowmap[r + 1]; ++k) {
          const int cc = g.colidx[k];
          if (cc < r0 || cc >= r) continue;
          for (int t = g.rowmap[cc]; t < g.rowmap[cc + 1]; ++t)
            if (g.colidx[t] == r) ch.tslot[k] = (uint16_t)(t - b0);
        }
    }
  });
  return ch;
}

}  // namespace pa
