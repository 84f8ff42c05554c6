// sd2d_host.hpp - host-side set-up of the device tables (no HIP calls):
// the per-cell face table, the tile-major metrics and the AV projections,
// built once from the arrays module setup2d holds after INIT_SETUP.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <map>
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "../../include/sd2d.h"
#include "sd2d_device.hpp"

namespace sd2d {

struct HostTables {
  int N = 0, ncell = 0, cpw = 0, ntile = 0;
  std::vector<int> c2s;      // cell -> slot (tile*cpw + position in the tile)
  std::vector<int> s2c;      // slot -> cell, -1 for a padding slot
  Ops ops{};
  Phys ph{};
  std::vector<int> ft;       // [tile][k][{neighbour base, info}][ci]
  std::vector<int> tvalid;   // valid cells of each tile (its first slots)
  std::vector<double> met;   // [tile][metric slot][ci]
  std::vector<double> av;    // AV smoothing filter F[is][k] (N x N)
  std::vector<double> metw;  // viscous path: other-family metrics at line ends
};

namespace hostdetail {

struct HostMesh {
  int N, nc, nface;
  const sd2d_mesh *m;
  int if2c(int f, int k) const { return m->if2c[f + (size_t)nface * k]; }
  // own metric normal of local face k (0-based) at own point nfp (1-based):
  // S1(1,:) on faces 2/4, S2(2,:) on faces 1/3 (interfaceflux_all.f90:47,54)
  void normal(int cell, int k, int nfp, double &nx, double &ny) const {
    int ifp, jfp;  // 1-based, face2fpconnec (initsetup.f90:356-380)
    switch (k) {
      case 0: ifp = nfp; jfp = 1; break;
      case 1: ifp = N + 1; jfp = nfp; break;
      case 2: ifp = N - nfp + 1; jfp = N + 1; break;
      default: ifp = 1; jfp = N - nfp + 1; break;
    }
    if (k == 1 || k == 3) {
      const size_t b = 2 * (2 * ((ifp - 1) + (size_t)(N + 1) * ((jfp - 1) + (size_t)N * cell)));
      nx = m->s1[b + 0]; ny = m->s1[b + 2];
    } else {
      const size_t b = 2 * (2 * ((ifp - 1) + (size_t)N * ((jfp - 1) + (size_t)(N + 1) * cell)));
      nx = m->s2[b + 1]; ny = m->s2[b + 3];
    }
  }
};

template <typename Tm>
inline void put_tile(std::vector<Tm> &dst, int NM, int CPW, int cell, int slot, Tm v) {
  const size_t tile = cell / CPW, ci = cell % CPW;
  dst[(tile * NM + slot) * CPW + ci] = v;
}

inline double lagrange(const double *x, int n, int i, double t) {
  double a = 1.0, b = 1.0;
  for (int s = 0; s < n; ++s)
    if (s != i) { a *= (t - x[s]); b *= (x[i] - x[s]); }
  return a / b;
}

}  // namespace hostdetail

// A face point whose right cell takes the left cell's normal (times the
// face's sign) in the device metric table.
struct NormalAdopt {
  int cell, k, nfp;   // right cell (0-based), its local face (0-based), own point (1-based)
  double nx, ny;
};

// Two normals of one face point that agree to this fraction of their
// length differ only by rounding: both cells map the shared edge through
// the same nodes (an 8-node edge through its three nodes, whose
// derivatives the two cells form in different orders).  A non-conforming
// pairing differs at O(1).
constexpr double NORMAL_RTOL = 1e-9;

// Per-cell face table.  Each cell evaluates its faces itself, so the two
// cells of an interior or periodic face must use one normal up to sign:
// the reference forms the face flux with the left cell's normal only
// (interfaceflux_all.f90:47,54, cyclocflux_all.f90:47,54).  Where the
// right cell's own vector equals -+ the left's bit for bit (bilinear
// cells) nothing changes; where they differ by rounding (curved 8-node
// cells) the face point is listed in `adopt`, and build_host_tables puts
// the left's normal into the right cell's metric slot.  A pair that
// differs beyond NORMAL_RTOL is refused.
inline int build_face_table(const sd2d_mesh *m, int CPW, const std::vector<int> &c2s,
                            std::vector<int> &nb, std::vector<int> &info,
                            std::vector<NormalAdopt> &adopt, std::string &err) {
  using hostdetail::HostMesh;
  const int N = m->N, nc = m->ncell, N2 = N * N;
  HostMesh hm{N, nc, m->nface, m};
  nb.assign((size_t)nc * 4, 0);
  info.assign((size_t)nc * 4, -1);
  adopt.clear();
  auto chk_face = [&](int f) { return f >= 1 && f <= m->nface; };
  auto base_of = [&](int cell) {
    const size_t slot = c2s[cell], tile = slot / CPW, ci = slot % CPW;
    return (long long)(tile * 9 * N2 * CPW + ci);
  };
  auto bad = [&](int code, const char *msg) { err = msg; return code; };
  auto pair = [&](int cl, int kl, int cr, int kr) -> int {
    // kl, kr 1-based local faces; cl, cr 0-based cells
    if (cl < 0 || cl >= nc || cr < 0 || cr >= nc || kl < 1 || kl > 4 || kr < 1 || kr > 4)
      return bad(SD2D_EINVAL, "face table entry out of range");
    int sneg = -1;
    for (int nfp = 1; nfp <= N; ++nfp) {
      double lx, ly, rx, ry;
      hm.normal(cl, kl - 1, nfp, lx, ly);
      hm.normal(cr, kr - 1, N - nfp + 1, rx, ry);
      int s;
      if (lx == rx && ly == ry) s = 0;
      else if (lx == -rx && ly == -ry) s = 1;
      else {
        const double tol = NORMAL_RTOL * std::max(std::hypot(lx, ly), std::hypot(rx, ry));
        if (std::fabs(lx - rx) <= tol && std::fabs(ly - ry) <= tol) s = 0;
        else if (std::fabs(lx + rx) <= tol && std::fabs(ly + ry) <= tol) s = 1;
        else
          return bad(SD2D_EUNSUPPORTED,
                     "face normals of neighbouring cells differ beyond sign and rounding "
                     "(non-conforming faces are not supported)");
        adopt.push_back({cr, kr - 1, N - nfp + 1, s ? -lx : lx, s ? -ly : ly});
      }
      if (sneg >= 0 && s != sneg)
        return bad(SD2D_EUNSUPPORTED, "face normal sign varies along a face");
      sneg = s;
    }
    const long long bl = base_of(cl), br = base_of(cr);
    if (br > 0x7fffffffLL || bl > 0x7fffffffLL)
      return bad(SD2D_EUNSUPPORTED, "mesh too large for 32-bit offsets");
    nb[4 * (size_t)cl + (kl - 1)] = (int)br;
    info[4 * (size_t)cl + (kl - 1)] = finfo_pack(kr - 1, 1, FT_INTERIOR, 0);
    nb[4 * (size_t)cr + (kr - 1)] = (int)bl;
    info[4 * (size_t)cr + (kr - 1)] = finfo_pack(kl - 1, 0, FT_INTERIOR, sneg);
    return SD2D_OK;
  };
  for (int i = 0; i < m->ninter; ++i) {
    const int f = m->ibfinter[i];
    if (!chk_face(f)) return bad(SD2D_EINVAL, "IBFINTER entry out of range");
    const int r = pair(hm.if2c(f - 1, 0) - 1, hm.if2c(f - 1, 2), hm.if2c(f - 1, 1) - 1,
                       hm.if2c(f - 1, 3));
    if (r) return r;
  }
  for (int i = 0; i < m->npair; ++i) {   // cyclocflux_all.f90:15-24
    const int k1 = m->cyc2fpair[i], k2 = m->cyc2fpair[i + m->npair];
    if (!chk_face(k1) || !chk_face(k2)) return bad(SD2D_EINVAL, "CYC2FPAIR out of range");
    const int r = pair(hm.if2c(k1 - 1, 0) - 1, hm.if2c(k1 - 1, 2), hm.if2c(k2 - 1, 0) - 1,
                       hm.if2c(k2 - 1, 2));
    if (r) return r;
  }
  // boundary lists in BCFLUX order (bcflux.f90: INLE, OUTL, WALL, SYMP, FREE)
  struct L { int n; const int32_t *f; int bc; } lists[5] = {
      {m->ninlet, m->ibfinl, FT_INLE}, {m->noutlet, m->ibfout, FT_OUTL},
      {m->nwall, m->ibfwal, FT_WALL}, {m->nsymp, m->ibfsymp, FT_SYMP},
      {m->nfree, m->ibfree, FT_FREE}};
  for (auto &l : lists)
    for (int i = 0; i < l.n; ++i) {
      const int f = l.f[i];
      if (!chk_face(f)) return bad(SD2D_EINVAL, "boundary list entry out of range");
      const int cl = hm.if2c(f - 1, 0) - 1, kl = hm.if2c(f - 1, 2);
      if (cl < 0 || cl >= nc || kl < 1 || kl > 4)
        return bad(SD2D_EINVAL, "boundary face table entry out of range");
      nb[4 * (size_t)cl + (kl - 1)] = 0;
      info[4 * (size_t)cl + (kl - 1)] = finfo_pack(0, 1, l.bc, 0);
    }
  for (size_t i = 0; i < info.size(); ++i)
    if (info[i] < 0)
      return bad(SD2D_EINVAL,
                 "a cell face is neither interior, periodic nor on a boundary list "
                 "(it would be a processor interface)");
  return SD2D_OK;
}

// Breadth-first recursive bisection of a graph given in CSR form.  A part
// [b, e) of the vertex order holding more than one grain of vertices is
// sorted by breadth-first distance from a pseudo-peripheral vertex and
// split after floor(grains/2) grains (the part's last grain takes any
// remainder), down to single grains; parts at every level are compact and
// consecutive vertices of the result lie close together in the graph.
inline std::vector<int> bisection_order(int n, int grain, const std::vector<int> &xadj,
                                        const std::vector<int> &ladj) {
  std::vector<int> ord(n), part(n, -1), dist(n, -1), queue;
  for (int t = 0; t < n; ++t) ord[t] = t;
  queue.reserve(n);
  auto bfs = [&](int pid, int src, int b, int e) {
    for (int i = b; i < e; ++i) dist[ord[i]] = -1;
    queue.clear();
    queue.push_back(src);
    dist[src] = 0;
    for (size_t h = 0; h < queue.size(); ++h) {
      const int t = queue[h];
      for (int j = xadj[t]; j < xadj[t + 1]; ++j) {
        const int u = ladj[j];
        if (part[u] == pid && dist[u] < 0) { dist[u] = dist[t] + 1; queue.push_back(u); }
      }
    }
    return queue.back();
  };
  std::vector<std::pair<int, int>> stack{{0, n}};
  int pid = 0;
  while (!stack.empty()) {
    const auto [b, e] = stack.back();
    stack.pop_back();
    const int ng = (e - b + grain - 1) / grain;   // grains in the part
    if (ng <= 1) continue;
    ++pid;
    for (int i = b; i < e; ++i) part[ord[i]] = pid;
    // pseudo-peripheral start: the farthest vertex from the lowest-numbered one
    const int first = *std::min_element(ord.begin() + b, ord.begin() + e);
    const int p = bfs(pid, first, b, e);
    bfs(pid, p, b, e);
    std::stable_sort(ord.begin() + b, ord.begin() + e, [&](int x, int y) {
      const long long dx = dist[x] < 0 ? (1LL << 40) : dist[x];
      const long long dy = dist[y] < 0 ? (1LL << 40) : dist[y];
      return dx != dy ? dx < dy : x < y;
    });
    const int m = b + (ng / 2) * grain;
    stack.push_back({m, e});     // second half after the first
    stack.push_back({b, m});
  }
  return ord;
}

// Row length of a mesh numbered row by row: R when every interior face and
// periodic pair joins cells c and c+1 of one row, c and c+R (next row), the
// first and last cells of a row (c and c+R-1, a periodic row) or a cell of
// the first and of the last row (c and c+nc-R, periodic columns); 0
// otherwise.  R is the most frequent distance above 1.
inline int mesh_row_length(int nc, const std::vector<std::pair<int, int>> &pairs) {
  std::map<int, long long> hist;
  for (const auto &pr : pairs) {
    const int d = std::abs(pr.first - pr.second);
    if (d > 1) ++hist[d];
  }
  int R = 0;
  long long best = 0;
  for (const auto &h : hist)
    if (h.second > best) { best = h.second; R = h.first; }
  if (R < 2 || nc % R != 0) return 0;
  for (const auto &pr : pairs) {
    const int lo = std::min(pr.first, pr.second), d = std::abs(pr.first - pr.second);
    const bool ok = (d == 1 && lo % R != R - 1) || d == R || (d == R - 1 && lo % R == 0) ||
                    (d == nc - R && lo < R);
    if (!ok) return 0;
  }
  return R;
}

// Cell -> slot map of a tile shape.  Strips: slot = cell, except that a
// mesh numbered row by row with rows of R >= CPW cells, R not a multiple of
// CPW, gets every row padded to a multiple of CPW (slot = row*Rpad + column)
// when that adds at most 1/16 of the row
// so that the tiles of one row line up with those of the next and a tile's
// neighbours across the rows lie in one tile.  Blocks: the cells in the
// order of a recursive bisection of the cell adjacency graph (interior
// faces and periodic pairs) with one tile as the grain, so every tile is a
// compact patch of cells.  Either way the padding slots of a tile come
// after its cells; s2c is -1 there.
inline void tile_slots(const sd2d_mesh *m, int CPW, int shape, std::vector<int> &c2s,
                       std::vector<int> &s2c) {
  const int nc = m->ncell;
  std::vector<std::pair<int, int>> pairs;   // 0-based cells joined by a face
  auto link = [&](int f1, int k1, int f2, int k2) {
    if (f1 < 1 || f1 > m->nface || f2 < 1 || f2 > m->nface) return;
    const int a = m->if2c[(f1 - 1) + (size_t)m->nface * k1] - 1;
    const int b = m->if2c[(f2 - 1) + (size_t)m->nface * k2] - 1;
    if (a >= 0 && a < nc && b >= 0 && b < nc && a != b) pairs.push_back({a, b});
  };
  for (int i = 0; i < m->ninter; ++i) link(m->ibfinter[i], 0, m->ibfinter[i], 1);
  for (int i = 0; i < m->npair; ++i) link(m->cyc2fpair[i], 0, m->cyc2fpair[i + m->npair], 0);
  std::vector<int> order(nc);
  for (int c = 0; c < nc; ++c) order[c] = c;
  c2s.assign(nc, 0);
  if (shape == SD2D_TILE_ROWPAIRS) {
    // two consecutive rows of CPW/2 cells each: the faces between the two
    // rows lie inside the tile, so every line across the rows has at most
    // one end in another tile.  A last odd row is cut into strips of CPW.
    const int R = mesh_row_length(nc, pairs), H = CPW / 2;
    if (R >= H) {
      const int nrow = nc / R, nch = (R + H - 1) / H, nfull = (R + CPW - 1) / CPW;
      const int nt = (nrow / 2) * nch + (nrow % 2) * nfull;
      s2c.assign((size_t)nt * CPW, -1);
      int t = 0;
      for (int k = 0; k + 1 < nrow; k += 2)
        for (int j = 0; j < nch; ++j, ++t) {
          const int c0 = j * H, w = std::min(H, R - c0);
          for (int i = 0; i < w; ++i) {
            const int a = k * R + c0 + i, b = (k + 1) * R + c0 + i;
            c2s[a] = t * CPW + i;
            c2s[b] = t * CPW + w + i;
            s2c[c2s[a]] = a;
            s2c[c2s[b]] = b;
          }
        }
      if (nrow % 2)
        for (int j = 0; j < nfull; ++j, ++t) {
          const int c0 = j * CPW, w = std::min(CPW, R - c0);
          for (int i = 0; i < w; ++i) {
            const int a = (nrow - 1) * R + c0 + i;
            c2s[a] = t * CPW + i;
            s2c[c2s[a]] = a;
          }
        }
      return;
    }
    shape = SD2D_TILE_STRIPS;   // not numbered row by row
  }
  if (shape == SD2D_TILE_BLOCKS) {
    std::vector<std::vector<int>> adj(nc);
    for (const auto &pr : pairs) {
      adj[pr.first].push_back(pr.second);
      adj[pr.second].push_back(pr.first);
    }
    std::vector<int> xadj(nc + 1, 0), ladj;
    for (int c = 0; c < nc; ++c) {
      std::sort(adj[c].begin(), adj[c].end());
      adj[c].erase(std::unique(adj[c].begin(), adj[c].end()), adj[c].end());
      xadj[c + 1] = xadj[c] + (int)adj[c].size();
      ladj.insert(ladj.end(), adj[c].begin(), adj[c].end());
    }
    order = bisection_order(nc, CPW, xadj, ladj);
  } else {
    const int R = mesh_row_length(nc, pairs);
    const int Rpad = R > 0 ? (R + CPW - 1) / CPW * CPW : 0;
    if (R >= CPW && Rpad != R && 16 * (Rpad - R) <= R) {   // padding at most 1/16
      s2c.assign((size_t)(nc / R) * Rpad, -1);
      for (int c = 0; c < nc; ++c) {
        c2s[c] = (c / R) * Rpad + c % R;
        s2c[c2s[c]] = c;
      }
      return;
    }
  }
  const int nt = (nc + CPW - 1) / CPW;
  s2c.assign((size_t)nt * CPW, -1);
  for (int sl = 0; sl < nc; ++sl) {
    c2s[order[sl]] = sl;
    s2c[sl] = order[sl];
  }
}

// Launch order of the tiles: recursive bisection of the tile adjacency
// graph (two tiles are adjacent when a face joins their cells).  Each part
// is split at the median of the breadth-first distance from a
// pseudo-peripheral tile of the part, so that every part, down to groups
// of a few tiles, is a compact patch of the mesh and consecutive tiles in
// the order are neighbours.  Tiles that are processed close together in
// time then find each other's states in the L2 cache of their XCD,
// whatever the cell numbering of the mesh.
inline std::vector<int> tile_order(int nt, int CPW, int tile_doubles,
                                   const std::vector<int> &nb, const std::vector<int> &info,
                                   const std::vector<int> &c2s) {
  // adjacency (CSR, deduplicated)
  std::vector<std::vector<int>> adj(nt);
  const int nc = (int)(nb.size() / 4);
  for (int cell = 0; cell < nc; ++cell)
    for (int k = 0; k < 4; ++k) {
      const size_t i = 4 * (size_t)cell + k;
      if (((info[i] >> 3) & 7) != FT_INTERIOR) continue;
      const int t = c2s[cell] / CPW, u = nb[i] / tile_doubles;
      if (u != t) adj[t].push_back(u);
    }
  std::vector<int> xadj(nt + 1, 0), ladj;
  for (int t = 0; t < nt; ++t) {
    std::sort(adj[t].begin(), adj[t].end());
    adj[t].erase(std::unique(adj[t].begin(), adj[t].end()), adj[t].end());
    xadj[t + 1] = xadj[t] + (int)adj[t].size();
    ladj.insert(ladj.end(), adj[t].begin(), adj[t].end());
  }
  constexpr int LEAF = 4;
  return bisection_order(nt, LEAF, xadj, ladj);
}

// Everything the kernels read besides the state.
inline int build_host_tables(const sd2d_mesh *m, const sd2d_params *p, HostTables &t,
                             std::string &err) {
  using hostdetail::put_tile;
  const int N = m->N, nc = m->ncell, CPW = cpw_of(N), N2 = N * N;
  t.N = N; t.ncell = nc; t.cpw = CPW; t.ntile = (nc + CPW - 1) / CPW;
  for (int i = 0; i <= N; ++i)
    for (int s = 0; s < N; ++s) {
      t.ops.L[i][s] = m->lmat[i + (N + 1) * s];
      t.ops.M[i][s] = m->mmat[i + (N + 1) * s];
    }
  t.ph.alpha = p->alpha; t.ph.rinf = p->rinf; t.ph.pinf = p->pinf; t.ph.uinf = p->uinf;
  t.ph.vinf = p->vinf; t.ph.bxinf = p->bxinf; t.ph.byinf = p->byinf;
  t.ph.vismode = p->vismode;

  if (p->tile_shape < SD2D_TILE_DEFAULT || p->tile_shape > SD2D_TILE_ROWPAIRS) {
    err = "tile_shape must be SD2D_TILE_DEFAULT, SD2D_TILE_STRIPS, SD2D_TILE_BLOCKS or "
          "SD2D_TILE_ROWPAIRS";
    return SD2D_EINVAL;
  }
  tile_slots(m, CPW, p->tile_shape == SD2D_TILE_DEFAULT ? SD2D_TILE_ROWPAIRS : p->tile_shape,
             t.c2s, t.s2c);
  t.ntile = (int)(t.s2c.size() / CPW);
  std::vector<int> nb, info;
  std::vector<NormalAdopt> adopt;
  int r = build_face_table(m, CPW, t.c2s, nb, info, adopt, err);
  if (r) return r;
  const int nt = t.ntile;
  // the tiles are renumbered in launch order, so that a workgroup finds its
  // tile from its position alone (no table read ahead of every other load);
  // a tile keeps the positions of its cells, so its valid cells stay a
  // prefix of it (tvalid)
  {
    std::vector<int> ord = tile_order(nt, CPW, 9 * N2 * CPW, nb, info, t.c2s);
    std::vector<int> pos(nt);
    for (int i = 0; i < nt; ++i) pos[ord[i]] = i;
    t.s2c.assign((size_t)nt * CPW, -1);
    for (int cell = 0; cell < nc; ++cell) {
      const int sl = t.c2s[cell];
      t.c2s[cell] = pos[sl / CPW] * CPW + sl % CPW;
      t.s2c[t.c2s[cell]] = cell;
    }
    nb.clear();
    info.clear();
    r = build_face_table(m, CPW, t.c2s, nb, info, adopt, err);
    if (r) return r;
  }
  t.ft.assign((size_t)nt * 8 * CPW, 0);
  for (int cell = 0; cell < nc; ++cell)
    for (int k = 0; k < 4; ++k) {
      put_tile<int>(t.ft, 8, CPW, t.c2s[cell], 2 * k, nb[4 * (size_t)cell + k]);
      put_tile<int>(t.ft, 8, CPW, t.c2s[cell], 2 * k + 1, info[4 * (size_t)cell + k]);
    }
  // padded slots: FREE boundary so that nothing is dereferenced for them
  for (int sl = 0; sl < nt * CPW; ++sl)
    if (t.s2c[sl] < 0)
      for (int k = 0; k < 4; ++k)
        put_tile<int>(t.ft, 8, CPW, sl, 2 * k + 1, finfo_pack(0, 1, FT_FREE, 0));
  // valid cells per tile; the kernels take a tile's cells as the first
  // tvalid[t] slots of it
  t.tvalid.assign(nt, 0);
  for (int tt = 0; tt < nt; ++tt) {
    int nv = 0;
    while (nv < CPW && t.s2c[(size_t)tt * CPW + nv] >= 0) ++nv;
    for (int ci = nv; ci < CPW; ++ci)
      if (t.s2c[(size_t)tt * CPW + ci] >= 0) {
        err = "internal error: the cells of a tile are not a prefix of it";
        return SD2D_EINVAL;
      }
    t.tvalid[tt] = nv;
  }

  // metrics: S1(1,1:2,rfp,js), S2(2,1:2,is,rfp), 1/Jac(is,js)
  const int NM = 4 * N * (N + 1) + N2;
  t.met.assign((size_t)nt * NM * CPW, 0.0);
  for (int cell = 0; cell < nc; ++cell) {
    for (int js = 0; js < N; ++js)
      for (int rfp = 0; rfp <= N; ++rfp) {
        const size_t b = 2 * (2 * (rfp + (size_t)(N + 1) * (js + (size_t)N * cell)));
        put_tile<double>(t.met, NM, CPW, t.c2s[cell], 0 + rfp + (N + 1) * js, m->s1[b + 0]);
        put_tile<double>(t.met, NM, CPW, t.c2s[cell], N * (N + 1) + rfp + (N + 1) * js, m->s1[b + 2]);
      }
    for (int rfp = 0; rfp <= N; ++rfp)
      for (int is = 0; is < N; ++is) {
        const size_t b = 2 * (2 * (is + (size_t)N * (rfp + (size_t)(N + 1) * cell)));
        put_tile<double>(t.met, NM, CPW, t.c2s[cell], 2 * N * (N + 1) + rfp + (N + 1) * is, m->s2[b + 1]);
        put_tile<double>(t.met, NM, CPW, t.c2s[cell], 3 * N * (N + 1) + rfp + (N + 1) * is, m->s2[b + 3]);
      }
    for (int js = 0; js < N; ++js)
      for (int is = 0; is < N; ++is)
        put_tile<double>(t.met, NM, CPW, t.c2s[cell], 4 * N * (N + 1) + js + N * is,
                         1.0 / m->jac[is + N * (js + (size_t)N * cell)]);
  }
  // the right cell of a face whose normals differ by rounding takes the
  // left cell's normal at that face point (face2fpconnec, as in
  // HostMesh::normal); everything else of the right cell keeps its own
  // metrics, so its divB/Jz weights at that point and its viscous terms use
  // a vector within rounding of the reference's own
  for (const NormalAdopt &a : adopt) {
    int ifp, jfp;
    switch (a.k) {
      case 0: ifp = a.nfp; jfp = 1; break;
      case 1: ifp = N + 1; jfp = a.nfp; break;
      case 2: ifp = N - a.nfp + 1; jfp = N + 1; break;
      default: ifp = 1; jfp = N - a.nfp + 1; break;
    }
    const int sl = t.c2s[a.cell];
    if (a.k == 1 || a.k == 3) {
      const int rfp = ifp - 1, js = jfp - 1;
      put_tile<double>(t.met, NM, CPW, sl, 0 + rfp + (N + 1) * js, a.nx);
      put_tile<double>(t.met, NM, CPW, sl, N * (N + 1) + rfp + (N + 1) * js, a.ny);
    } else {
      const int rfp = jfp - 1, is = ifp - 1;
      put_tile<double>(t.met, NM, CPW, sl, 2 * N * (N + 1) + rfp + (N + 1) * is, a.nx);
      put_tile<double>(t.met, NM, CPW, sl, 3 * N * (N + 1) + rfp + (N + 1) * is, a.ny);
    }
  }
  // viscous path: S1(2,1:2,rfp,js) and S2(1,1:2,is,rfp) at the two ends of
  // every line, used by the adiabatic-wall gradient ghost (bcflux.f90:547-571)
  if (p->vismode == 1 && (p->mu != 0.0 || p->eta0 != 0.0)) {
    const int NW = 8 * N;
    t.metw.assign((size_t)nt * NW * CPW, 0.0);
    for (int cell = 0; cell < nc; ++cell)
      for (int end = 0; end < 2; ++end) {
        const int rfp = end ? N : 0;
        for (int line = 0; line < N; ++line) {
          const size_t b1 = 2 * (2 * (rfp + (size_t)(N + 1) * (line + (size_t)N * cell)));
          const size_t b2 = 2 * (2 * (line + (size_t)N * (rfp + (size_t)(N + 1) * cell)));
          for (int comp = 0; comp < 2; ++comp) {
            put_tile<double>(t.metw, NW, CPW, t.c2s[cell], 0 * 4 * N + end * 2 * N + comp * N + line,
                             m->s1[b1 + 1 + 2 * comp]);
            put_tile<double>(t.metw, NW, CPW, t.c2s[cell], 1 * 4 * N + end * 2 * N + comp * N + line,
                             m->s2[b2 + 0 + 2 * comp]);
          }
        }
      }
  }
  // AV smoothing filter per direction (AV.f90:18-63): the projection onto
  // the N-1 low-order points, hlow[isl][is] = h_is(Xs_low(isl)), followed by
  // the interpolation back, hhigh[is][isl] = hlow_isl(Xs(is)), as one map
  // F[is][k] = sum_isl hhigh[is][isl] hlow[isl][k]
  t.av.assign((size_t)N * N, 0.0);
  if (m->xs_low) {
    std::vector<double> hlow((size_t)(N - 1) * N), hhigh((size_t)N * (N - 1));
    for (int il = 0; il < N - 1; ++il)
      for (int is = 0; is < N; ++is)
        hlow[il * N + is] = hostdetail::lagrange(m->xs, N, is, m->xs_low[il]);
    for (int is = 0; is < N; ++is)
      for (int il = 0; il < N - 1; ++il)
        hhigh[is * (N - 1) + il] = hostdetail::lagrange(m->xs_low, N - 1, il, m->xs[is]);
    for (int is = 0; is < N; ++is)
      for (int k = 0; k < N; ++k) {
        double f = 0.0;
        for (int il = 0; il < N - 1; ++il) f += hhigh[is * (N - 1) + il] * hlow[il * N + k];
        t.av[is * N + k] = f;
      }
  }
  return SD2D_OK;
}

}  // namespace sd2d
