// Geodesic-distance adjacency (host): every pair of vertices of a triangle
// mesh closer than a distance along the surface, the sets adjacency.pyx:44-79
// (compute) and :82-106 (compute_distance_parallel) build with an exact
// geodesic algorithm, once per mesh before any permutation.
//
// Method: exact distances by window propagation (the continuous Dijkstra of
// Mitchell, Mount and Papadimitriou, in the edge-window form of Surazhsky et
// al. 2005), written here from the method's definition.  A window is an
// interval of an edge whose points all see one source image, unfolded into the
// plane of the face the window enters, at the distance offset of its
// (pseudo-)source.  A window crossing a face splits at the face's third vertex
// into windows on the two other edges; geodesics bend only at saddle or
// boundary vertices, which become pseudo-sources at their own distance.  No
// window is trimmed against another (the search radius is a few edge
// lengths, where windows stay few), so the distance of a vertex is the least
// over every window that reaches it, and the result is the true surface
// distance up to rounding.  Windows are taken in order of their least
// distance; those farther than the radius are dropped.  Sources are
// independent and run on host threads.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <limits>
#include <queue>
#include <string>
#include <thread>
#include <vector>

#include "capi_common.h"

using tfcehip::set_error;

namespace {

constexpr double kInf = std::numeric_limits<double>::infinity();

struct Geo {
  int nv = 0, nf = 0, ne = 0;
  std::vector<double> p;        // [nv][3]
  std::vector<int32_t> fv;      // [nf][3]
  std::vector<int32_t> ev;      // [ne][2], ev[2e] < ev[2e+1]
  std::vector<int32_t> ef;      // [ne][2] faces on the edge (-1: none)
  std::vector<double> el;       // [ne] lengths
  std::vector<int32_t> fe;      // [nf][3] edge opposite corner k of the face
  std::vector<int32_t> vf_ptr;  // faces around each vertex (CSR)
  std::vector<int32_t> vf;
  std::vector<int32_t> vn_ptr;  // neighbours of each vertex (CSR)
  std::vector<int32_t> vn;
  std::vector<char> bend;       // saddle or boundary vertex: a geodesic may bend there
};

double dist3(const Geo& g, int a, int b) {
  const double dx = g.p[3 * a] - g.p[3 * b], dy = g.p[3 * a + 1] - g.p[3 * b + 1], dz = g.p[3 * a + 2] - g.p[3 * b + 2];
  return std::sqrt(dx * dx + dy * dy + dz * dz);
}

// Position of vertex c in the frame of edge e (endpoint ev[2e] at the origin,
// ev[2e+1] on the +x axis), on the +y side when up, else on the -y side.
void place(const Geo& g, int e, int c, bool up, double& x, double& y) {
  const double L = g.el[e];
  const double a = dist3(g, g.ev[2 * e], c), b = dist3(g, g.ev[2 * e + 1], c);
  x = (a * a - b * b + L * L) / (2.0 * L);
  const double h2 = a * a - x * x;
  y = h2 > 0.0 ? std::sqrt(h2) : 0.0;
  if (!up) y = -y;
}

int third(const Geo& g, int f, int e) {
  for (int k = 0; k < 3; ++k)
    if (g.fe[3 * f + k] == e) return g.fv[3 * f + k];
  return -1;
}

int build(Geo& g, const float* v, int32_t nv, const int32_t* f, int32_t nf) {
  g.nv = nv;
  g.nf = nf;
  g.p.assign(v, v + 3 * (size_t)nv);
  g.fv.assign(f, f + 3 * (size_t)nf);
  for (int32_t i = 0; i < 3 * nf; ++i)
    if (g.fv[i] < 0 || g.fv[i] >= nv) return set_error(TFCE_EADJ, "a face names a vertex out of range");
  // edges: sort the 3 nf (min, max, face, corner) records
  struct Rec {
    int32_t a, b, f, k;
  };
  std::vector<Rec> r;
  r.reserve(3 * (size_t)nf);
  for (int32_t i = 0; i < nf; ++i)
    for (int k = 0; k < 3; ++k) {
      const int32_t a = g.fv[3 * i + (k + 1) % 3], b = g.fv[3 * i + (k + 2) % 3];  // opposite corner k
      r.push_back({std::min(a, b), std::max(a, b), i, k});
    }
  std::sort(r.begin(), r.end(), [](const Rec& x, const Rec& y) { return x.a != y.a ? x.a < y.a : x.b < y.b; });
  g.fe.assign(3 * (size_t)nf, -1);
  for (size_t i = 0; i < r.size();) {
    size_t j = i;
    while (j < r.size() && r[j].a == r[i].a && r[j].b == r[i].b) ++j;
    const int e = g.ne++;
    g.ev.push_back(r[i].a);
    g.ev.push_back(r[i].b);
    g.ef.push_back(r[i].f);
    g.ef.push_back(j - i > 1 ? r[i + 1].f : -1);  // a third face on one edge is not followed
    g.el.push_back(dist3(g, r[i].a, r[i].b));
    for (size_t q = i; q < j; ++q) g.fe[3 * (size_t)r[q].f + r[q].k] = e;
    i = j;
  }
  // vertex -> faces, vertex -> neighbours
  g.vf_ptr.assign(nv + 1, 0);
  for (int32_t i = 0; i < 3 * nf; ++i) g.vf_ptr[g.fv[i] + 1]++;
  for (int32_t i = 0; i < nv; ++i) g.vf_ptr[i + 1] += g.vf_ptr[i];
  g.vf.resize(3 * (size_t)nf);
  {
    std::vector<int32_t> cur(g.vf_ptr.begin(), g.vf_ptr.end() - 1);
    for (int32_t i = 0; i < nf; ++i)
      for (int k = 0; k < 3; ++k) g.vf[cur[g.fv[3 * i + k]]++] = i;
  }
  g.vn_ptr.assign(nv + 1, 0);
  for (int e = 0; e < g.ne; ++e) {
    g.vn_ptr[g.ev[2 * e] + 1]++;
    g.vn_ptr[g.ev[2 * e + 1] + 1]++;
  }
  for (int32_t i = 0; i < nv; ++i) g.vn_ptr[i + 1] += g.vn_ptr[i];
  g.vn.resize(2 * (size_t)g.ne);
  {
    std::vector<int32_t> cur(g.vn_ptr.begin(), g.vn_ptr.end() - 1);
    for (int e = 0; e < g.ne; ++e) {
      g.vn[cur[g.ev[2 * e]]++] = g.ev[2 * e + 1];
      g.vn[cur[g.ev[2 * e + 1]]++] = g.ev[2 * e];
    }
  }
  // total angle and boundary status of every vertex; a vertex whose angle
  // reaches 2 pi (up to rounding) is taken as a bending vertex too, which can
  // only add valid paths
  std::vector<double> angle(nv, 0.0);
  g.bend.assign(nv, 0);
  for (int32_t i = 0; i < nf; ++i)
    for (int k = 0; k < 3; ++k) {
      const int a = g.fv[3 * i + k], b = g.fv[3 * i + (k + 1) % 3], c = g.fv[3 * i + (k + 2) % 3];
      const double ab = dist3(g, a, b), ac = dist3(g, a, c), bc = dist3(g, b, c);
      if (ab <= 0.0 || ac <= 0.0) continue;
      double cs = (ab * ab + ac * ac - bc * bc) / (2.0 * ab * ac);
      cs = std::max(-1.0, std::min(1.0, cs));
      angle[a] += std::acos(cs);
    }
  for (int e = 0; e < g.ne; ++e)
    if (g.ef[2 * e + 1] < 0) g.bend[g.ev[2 * e]] = g.bend[g.ev[2 * e + 1]] = 1;
  for (int32_t i = 0; i < nv; ++i)
    if (angle[i] >= 2.0 * M_PI - 1e-6) g.bend[i] = 1;
  return 0;
}

struct Win {
  int e, f;        // edge, and the face the window enters
  double b0, b1;   // interval along the edge from ev[2e]
  double sx, sy;   // source image in the edge's frame (the face on +y, the image at y <= 0)
  double sigma;    // distance at the (pseudo-)source
  double dmin;     // least distance of the interval
};
struct WinLater {
  bool operator()(const Win& a, const Win& b) const { return a.dmin > b.dmin; }
};

double seg_min(double sx, double sy, double b0, double b1) {
  const double x = std::min(std::max(sx, b0), b1);
  return std::sqrt((sx - x) * (sx - x) + sy * sy);
}

double cross(double ax, double ay, double bx, double by) { return ax * by - ay * bx; }

class Source {
 public:
  explicit Source(const Geo& g) : g_(g), d_(g.nv, kInf), spawned_(g.nv, kInf) {}

  // distances of every vertex within reach of s (d_ < reach), then touched_
  void run(int s, double reach) {
    for (int v : touched_) {
      d_[v] = kInf;
      spawned_[v] = kInf;
    }
    touched_.clear();
    q_ = std::priority_queue<Win, std::vector<Win>, WinLater>();
    reach_ = reach;
    update(s, 0.0);
    spawn(s, 0.0);
    while (!q_.empty()) {
      const Win w = q_.top();
      q_.pop();
      if (w.dmin >= reach_) break;
      step(w);
    }
  }
  const std::vector<int>& touched() const { return touched_; }
  double dist(int v) const { return d_[v]; }

 private:
  void update(int v, double d) {
    if (!(d < d_[v])) return;
    if (d_[v] == kInf && spawned_[v] == kInf) touched_.push_back(v);
    d_[v] = d;
  }

  // v as a source at distance sigma: its neighbours along edges, and one
  // window over the opposite edge of every face around it
  void spawn(int v, double sigma) {
    if (!(sigma < spawned_[v]) || sigma >= reach_) return;
    if (d_[v] == kInf && spawned_[v] == kInf) touched_.push_back(v);
    spawned_[v] = sigma;
    for (int k = g_.vn_ptr[v]; k < g_.vn_ptr[v + 1]; ++k) {
      const int u = g_.vn[k];
      set(u, sigma + dist3(g_, v, u));
    }
    for (int k = g_.vf_ptr[v]; k < g_.vf_ptr[v + 1]; ++k) {
      const int f = g_.vf[k];
      int e = -1;
      for (int c = 0; c < 3; ++c)
        if (g_.fv[3 * f + c] == v) e = g_.fe[3 * f + c];
      if (e < 0) continue;
      const int f2 = g_.ef[2 * e] == f ? g_.ef[2 * e + 1] : g_.ef[2 * e];
      if (f2 < 0) continue;  // boundary edge: its vertices are neighbours of v
      Win w;
      w.e = e;
      w.f = f2;
      w.b0 = 0.0;
      w.b1 = g_.el[e];
      place(g_, e, v, false, w.sx, w.sy);  // v lies in f, on the side opposite f2
      w.sigma = sigma;
      w.dmin = sigma + seg_min(w.sx, w.sy, w.b0, w.b1);
      if (w.dmin < reach_) q_.push(w);
    }
  }

  // a distance for v: bending vertices become pseudo-sources
  void set(int v, double d) {
    if (!(d < d_[v])) return;
    update(v, d);
    if (g_.bend[v]) spawn(v, d);
  }

  void step(const Win& w) {
    const double L = g_.el[w.e];
    const int v0 = g_.ev[2 * w.e], v1 = g_.ev[2 * w.e + 1];
    const double eps = 1e-12 * L;
    if (w.b0 <= eps) set(v0, w.sigma + std::sqrt(w.sx * w.sx + w.sy * w.sy));
    if (w.b1 >= L - eps) set(v1, w.sigma + std::sqrt((w.sx - L) * (w.sx - L) + w.sy * w.sy));
    // A window longer at every point of its interval than the path through
    // one end of the edge (that end's distance, then along the edge) only
    // carries paths that the end's route beats: every ray continues straight
    // from its point, and the route through the end to that point and on is
    // a surface path at most as long.  Along the edge, window minus route is
    // non-increasing for the route through v0 and non-decreasing for the one
    // through v1, so one interval end decides each.
    const double w1 = w.sigma + std::sqrt((w.b1 - w.sx) * (w.b1 - w.sx) + w.sy * w.sy);
    const double w0 = w.sigma + std::sqrt((w.b0 - w.sx) * (w.b0 - w.sx) + w.sy * w.sy);
    const double tol = 1e-12 * (w0 + w1);
    if (w1 > d_[v0] + w.b1 + tol || w0 > d_[v1] + (L - w.b0) + tol) return;
    const int c = third(g_, w.f, w.e);
    if (c < 0) return;
    double cx, cy;
    place(g_, w.e, c, true, cx, cy);
    if (cy <= eps || w.sy >= 0.0) return;  // flat face or a source on the edge's line
    // where the ray from the source image through c crosses the edge
    const double xc = w.sx + (cx - w.sx) * (-w.sy) / (cy - w.sy);
    if (xc >= w.b0 && xc <= w.b1)
      set(c, w.sigma + std::sqrt((cx - w.sx) * (cx - w.sx) + (cy - w.sy) * (cy - w.sy)));
    // rays through [b0, min(b1, xc)] leave through edge (v0, c), the rest
    // through edge (v1, c)
    if (w.b0 < xc) emit(w, edge_of(w.f, v0, c), v0, 0.0, 0.0, cx, cy, L, 0.0, w.b0, std::min(w.b1, xc));
    if (w.b1 > xc) emit(w, edge_of(w.f, v1, c), v1, L, 0.0, cx, cy, 0.0, 0.0, std::max(w.b0, xc), w.b1);
  }

  int edge_of(int f, int a, int b) const {
    for (int k = 0; k < 3; ++k) {
      const int e = g_.fe[3 * f + k];
      if ((g_.ev[2 * e] == a && g_.ev[2 * e + 1] == b) || (g_.ev[2 * e] == b && g_.ev[2 * e + 1] == a)) return e;
    }
    return -1;
  }

  // New window on edge e2 of the current face, from vertex a (at A=(ax,ay) in
  // the current frame, on the current edge) to the face's third vertex (at
  // B=(bx,by)), hit by the rays through [x0, x1] of the current edge; (ox, oy)
  // is the current face's vertex off e2, which lies on e2's source side.
  void emit(const Win& w, int e2, int a, double ax, double ay, double bx, double by, double ox, double oy, double x0,
            double x1) {
    if (e2 < 0 || x1 - x0 <= 1e-14 * g_.el[w.e]) return;
    const int f2 = g_.ef[2 * e2] == w.f ? g_.ef[2 * e2 + 1] : g_.ef[2 * e2];
    // hit points of the two bounding rays on the line A-B, as fractions s
    auto hit = [&](double x) {
      const double dx = x - w.sx, dy = -w.sy;
      const double den = cross(bx - ax, by - ay, dx, dy);
      double s = den != 0.0 ? cross(w.sx - ax, w.sy - ay, dx, dy) / den : 0.0;
      return std::min(1.0, std::max(0.0, s));
    };
    const double s0 = hit(x0), s1 = hit(x1);
    // frame of e2: origin at its lower-index endpoint
    const bool a_first = g_.ev[2 * e2] == a;
    double px, py, qx, qy;  // e2's first and second endpoints in the current frame
    if (a_first) {
      px = ax; py = ay; qx = bx; qy = by;
    } else {
      px = bx; py = by; qx = ax; qy = ay;
    }
    const double len = std::sqrt((qx - px) * (qx - px) + (qy - py) * (qy - py));
    if (len <= 0.0) return;
    const double ux = (qx - px) / len, uy = (qy - py) / len;
    double nx = -uy, ny = ux;
    if ((ox - px) * nx + (oy - py) * ny > 0.0) {
      nx = -nx;
      ny = -ny;
    }
    // positions along e2 of the hit points (from e2's first endpoint)
    const double h0x = ax + s0 * (bx - ax), h0y = ay + s0 * (by - ay);
    const double h1x = ax + s1 * (bx - ax), h1y = ay + s1 * (by - ay);
    double t0 = (h0x - px) * ux + (h0y - py) * uy, t1 = (h1x - px) * ux + (h1y - py) * uy;
    if (t0 > t1) std::swap(t0, t1);
    const double L2 = g_.el[e2];
    t0 = std::max(0.0, t0);
    t1 = std::min(L2, t1);
    if (t1 - t0 <= 1e-14 * L2) return;
    Win n;
    n.e = e2;
    n.f = f2;
    n.b0 = t0;
    n.b1 = t1;
    n.sx = (w.sx - px) * ux + (w.sy - py) * uy;
    n.sy = (w.sx - px) * nx + (w.sy - py) * ny;
    n.sigma = w.sigma;
    n.dmin = n.sigma + seg_min(n.sx, n.sy, n.b0, n.b1);
    if (n.dmin >= reach_) return;
    if (f2 < 0) {
      // boundary edge: nothing beyond, but its endpoints may be reached
      if (n.b0 <= 1e-12 * L2) set(g_.ev[2 * e2], n.sigma + std::sqrt(n.sx * n.sx + n.sy * n.sy));
      if (n.b1 >= L2 * (1.0 - 1e-12))
        set(g_.ev[2 * e2 + 1], n.sigma + std::sqrt((n.sx - L2) * (n.sx - L2) + n.sy * n.sy));
      return;
    }
    q_.push(n);
  }

  const Geo& g_;
  std::vector<double> d_, spawned_;
  std::vector<int> touched_;
  std::priority_queue<Win, std::vector<Win>, WinLater> q_;
  double reach_ = 0.0;
};

}  // namespace

struct tfce_geodesic {
  std::vector<int32_t> i, j;
  std::vector<double> d;
};

extern "C" {

int tfce_geodesic_pairs(const float* vertices, int32_t n_vertices, const int32_t* faces, int32_t n_faces,
                        double max_distance, int32_t source_begin, int32_t source_end, int32_t n_threads,
                        tfce_geodesic_t** out) {
  if (!vertices || !faces || !out) return set_error(TFCE_EINVAL, "NULL argument");
  if (n_vertices < 1 || n_faces < 1) return set_error(TFCE_EINVAL, "need at least one vertex and one face");
  if (!(max_distance > 0.0) || !std::isfinite(max_distance))
    return set_error(TFCE_EINVAL, "max_distance must be positive and finite");
  if (source_end <= 0) source_end = n_vertices;
  if (source_begin < 0 || source_begin > source_end || source_end > n_vertices)
    return set_error(TFCE_EINVAL, "source range out of [0, n_vertices]");
  Geo g;
  int rc = build(g, vertices, n_vertices, faces, n_faces);
  if (rc) return rc;
  int nt = n_threads > 0 ? n_threads : (int)std::max(1u, std::thread::hardware_concurrency());
  nt = std::min(nt, 256);
  std::atomic<int> next(source_begin);
  std::vector<tfce_geodesic> part(nt);
  auto worker = [&](int t) {
    Source src(g);
    constexpr int kChunk = 64;
    while (true) {
      const int s0 = next.fetch_add(kChunk);
      if (s0 >= source_end) break;
      for (int s = s0; s < std::min(source_end, s0 + kChunk); ++s) {
        src.run(s, max_distance);
        std::vector<int> near;
        for (int v : src.touched())
          if (v > s && src.dist(v) < max_distance) near.push_back(v);
        std::sort(near.begin(), near.end());
        for (int v : near) {
          part[t].i.push_back(s);
          part[t].j.push_back(v);
          part[t].d.push_back(src.dist(v));
        }
      }
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; ++t) pool.emplace_back(worker, t);
  worker(0);
  for (auto& th : pool) th.join();
  auto* res = new tfce_geodesic;
  size_t total = 0;
  for (auto& p : part) total += p.i.size();
  std::vector<size_t> order(total);
  res->i.reserve(total);
  res->j.reserve(total);
  res->d.reserve(total);
  for (auto& p : part) {
    res->i.insert(res->i.end(), p.i.begin(), p.i.end());
    res->j.insert(res->j.end(), p.j.begin(), p.j.end());
    res->d.insert(res->d.end(), p.d.begin(), p.d.end());
  }
  // pairs in (i, j) order, as the reference's loops produce them
  for (size_t k = 0; k < total; ++k) order[k] = k;
  std::sort(order.begin(), order.end(), [&](size_t a, size_t b) {
    return res->i[a] != res->i[b] ? res->i[a] < res->i[b] : res->j[a] < res->j[b];
  });
  tfce_geodesic* sorted = new tfce_geodesic;
  sorted->i.resize(total);
  sorted->j.resize(total);
  sorted->d.resize(total);
  for (size_t k = 0; k < total; ++k) {
    sorted->i[k] = res->i[order[k]];
    sorted->j[k] = res->j[order[k]];
    sorted->d[k] = res->d[order[k]];
  }
  delete res;
  *out = sorted;
  return 0;
}

int tfce_geodesic_count(const tfce_geodesic_t* h, int64_t* n_pairs) {
  if (!h || !n_pairs) return set_error(TFCE_EINVAL, "NULL argument");
  *n_pairs = (int64_t)h->i.size();
  return 0;
}

int tfce_geodesic_copy(const tfce_geodesic_t* h, int32_t* i, int32_t* j, double* d) {
  if (!h || !i || !j || !d) return set_error(TFCE_EINVAL, "NULL argument");
  std::copy(h->i.begin(), h->i.end(), i);
  std::copy(h->j.begin(), h->j.end(), j);
  std::copy(h->d.begin(), h->d.end(), d);
  return 0;
}

int tfce_geodesic_destroy(tfce_geodesic_t* h) {
  delete h;
  return 0;
}

}  // extern "C"
