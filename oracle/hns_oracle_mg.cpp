// CPU oracle of the multigrid-preconditioned pressure solve (TEST INFRASTRUCTURE ONLY: tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline load it; no product path does).
//
// The reference has no multigrid: it runs a fixed count of red-black Gauss-Seidel sweeps
// (src/Cuda/HNanoSolver.cu:364-373, redBlackGaussSeidelUpdate in src/Cuda/Kernel.cu:463-489). This
// file restates the V-cycle the library uses as the preconditioner of its to-tolerance solve, voxel
// by voxel on leaf-dense vectors, in the same FP32 operation order as the library's unfused
// per-operation kernels, so the library's V-cycle (every variant) can be checked bit for bit, and
// the flexible-CG solve on it (FP64 dot products) as the same algorithm on the host cores.
//
// Hierarchy (levels share the 8^3-leaf topology): level l = 0, 1, 2, 3 has S = 8 >> l cells per leaf
// axis; a level-l vector is leaf-dense, cell (leaf, x, y, z) at leaf * S^3 + (x * S + y) * S + z
// (at level 0 the payload order, x << 6 | y << 3 | z). Prolongation copies a coarse value to its 8
// children, restriction sums them; level l carries 4^l (6 u - face neighbours), zero outside the
// active leaves. Level 4 groups the leaves in 2^3-leaf blocks (block = leaf coordinate >> 1) with
// the Galerkin operator of level 3 under piecewise-constant prolongation: diagonal D_B = the sum over
// the block's leaves of 6 minus their face neighbours inside the block, and -n_BB' (the leaf face
// pairs between the blocks) towards each face-adjacent block. Red-black colours are the parity of
// the level coordinate. The shape: level 0 pre-smoothed by one pointwise step from zero,
// z' = 1.75 r * 0.166666667f, and post-smoothed by one red-black Gauss-Seidel iteration (relaxation
// 1.0, colour 1 first); levels 1-2 by one red-black SOR iteration at 1.5 on the way down and one
// (level 1) or two (level 2) on the way up, in the reverse colour order; every restricted residual
// scaled by 1.6, except level 2's by 1.8; level 3 by one SOR iteration at 1.1, the level-4
// correction (6 SOR iterations at 1.7 on the residual restricted by 1.6) and three iterations at 1.9
// in the reverse order; with more than 4096 blocks of a colour, 12 SOR iterations at 1.85 on
// level 3 alone.
//
// Parity pins (tests/test_mg_oracle.py): every level operator against the 7-point operator built by
// coordinate lookup, the Galerkin identities R A_l P = 4 A_{l+1} and the level-4 weights against
// P^T A_3 P (exact on integer vectors); the V-cycle is linear in r and contracts the energy-norm
// error of the Poisson problem below 0.5 per application; the flexible CG on it reaches the scipy
// solution, and the scene's projection to the FP64 oracle's residual.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <unordered_map>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace oracle_mg {

struct Shape {
  float jacobi0 = 1.75f, omega0 = 1.0f, omega_l = 1.5f, over = 1.6f, over_deep = 1.6f;
  int post2 = 2;       // red-black post-smoothing iterations on level 2 (one on level 1)
  float over3 = 1.8f;  // scale of the level-2 -> 3 restriction (over_deep: level 1 -> 2)
  int c3_pre = 1, c3_post = 3, c4_iters = 6;
  float c3_omega = 1.1f, c3_omega_post = 1.9f, c4_omega = 1.7f, c4_over = 1.6f;
  int coarse = 12;  // level 3 alone (too many blocks for level 4)
  float coarse_omega = 1.85f;
  int max_blocks_per_colour = 4096;
};

struct MG {
  int64_t L = 0;
  std::vector<std::array<int32_t, 3>> lc;  // leaf coordinates (origin >> 3)
  std::vector<std::array<int32_t, 6>> fn;  // face-neighbour leaves +x, -x, +y, -y, +z, -z (-1: absent)
  std::vector<int> lcol;                   // level-3 colour (leaf parity)
  // level 4
  bool two = false;
  int64_t B = 0;
  std::vector<int32_t> blk;                // block of each leaf
  std::vector<std::array<int32_t, 8>> child;  // q = 4 dx + 2 dy + dz (-1: absent)
  std::vector<std::array<int32_t, 6>> bnb;    // face-adjacent blocks
  std::vector<std::array<int, 6>> cnt;        // n_BB' per face
  std::vector<int> D;                         // D_B
  std::vector<float> invd;                    // 1 / D_B
  std::vector<int> bcol;
  // work vectors: levels 1..3 right-hand sides and solutions, level-0 residual, level-3 residual,
  // level 4
  std::vector<float> b[4], u[4], res0, rho3, b4, e4;
  Shape M;
};

// face q of a cell (x, y, z) of a level with S cells per leaf axis: the neighbouring cell's index in
// a leaf-dense level vector, or -1 when the neighbouring leaf is absent
inline int64_t face_cell(const MG& g, int S, int64_t leaf, int x, int y, int z, int q) {
  int c[3] = {x, y, z};
  const int ax = q >> 1, dir = (q & 1) ? -1 : 1;
  c[ax] += dir;
  int64_t nl = leaf;
  if (c[ax] < 0 || c[ax] >= S) {
    nl = g.fn[leaf][q];
    if (nl < 0) return -1;
    c[ax] = (c[ax] + S) % S;
  }
  return nl * S * S * S + ((int64_t)c[0] * S + c[1]) * S + c[2];
}
// the six neighbours' sum in the order +x, -x, +y, -y, +z, -z, absent ones adding 0
inline float nsum(const MG& g, int S, const float* u, int64_t leaf, int x, int y, int z) {
  float s = 0.f;
  for (int q = 0; q < 6; ++q) {
    const int64_t j = face_cell(g, S, leaf, x, y, z, q);
    const float v = j < 0 ? 0.f : u[j];
    s = q == 0 ? v : s + v;
  }
  return s;
}

// the red-black SOR step of the library's level-0 sweep (Kernel.cu:487-488's constant), with the
// Poisson right-hand side b and dx^2 = -1: pGS = (s + b) / 6
inline float sor0(float s, float b, float old, float omega) {
  const float pGS = (s - b * -1.0f) * 0.166666667f;
  return old + omega * (pGS - old);
}

// one level-0 colour sweep of A z = r (leaf-dense); first: z starts at zero
void sweep0(const MG& g, const float* r, float* z, int colour, bool first, float omega) {
#pragma omp parallel for schedule(static)
  for (int64_t l = 0; l < g.L; ++l)
    for (int v = 0; v < 512; ++v) {
      const int x = v >> 6, y = (v >> 3) & 7, zz = v & 7;
      if (((x + y + zz) & 1) != colour) continue;
      const int64_t i = l * 512 + v;
      const float s = (first && colour == 0) ? 0.f : nsum(g, 8, z, l, x, y, zz);
      z[i] = sor0(s, r[i], first ? 0.f : z[i], omega);
    }
}

// one colour sweep at level lgS >= 1 on scale (6 u - sum) = b
void sweep_l(const MG& g, int S, const float* b, float* u, float inv_s, int colour, bool first, float omega) {
  const int C = S * S * S;
#pragma omp parallel for schedule(static)
  for (int64_t l = 0; l < g.L; ++l)
    for (int c = 0; c < C; ++c) {
      const int x = c / (S * S), y = (c / S) % S, z = c % S;
      const int64_t i = l * C + c;
      if (((x + y + z) & 1) != colour) {
        if (first && colour == 0) u[i] = 0.f;
        continue;
      }
      const float s = (first && colour == 0) ? 0.f : nsum(g, S, u, l, x, y, z);
      const float gs = (b[i] * inv_s + s) * (1.0f / 6.0f);
      u[i] = first ? omega * gs : std::fma(omega, gs - u[i], u[i]);
    }
}

// residual of the level with S cells per axis, scale * (6 u - sum) against b, restricted (times
// over) to the next level: bc[leaf * (S/2)^3 + cell] = over * sum of the 8 children
void restrict_l(const MG& g, int S, const float* b, const float* u, float scale, float over, float* bc) {
  const int C = S * S * S, Sc = S / 2, Cc = Sc * Sc * Sc;
#pragma omp parallel for schedule(static)
  for (int64_t l = 0; l < g.L; ++l) {
    float res[512];
    for (int c = 0; c < C; ++c) {
      const int x = c / (S * S), y = (c / S) % S, z = c % S;
      const int64_t i = l * C + c;
      const float s = nsum(g, S, u, l, x, y, z);
      res[c] = S == 8 ? b[i] - (6.0f * u[i] - s) : b[i] - scale * (6.0f * u[i] - s);
    }
    for (int cc = 0; cc < Cc; ++cc) {
      const int X = cc / (Sc * Sc), Y = (cc / Sc) % Sc, Z = cc % Sc;
      float sum = 0.f;
      for (int k = 0; k < 8; ++k) {
        const int x = 2 * X + (k >> 2), y = 2 * Y + ((k >> 1) & 1), z = 2 * Z + (k & 1);
        sum += res[(x * S + y) * S + z];
      }
      bc[l * Cc + cc] = over * sum;
    }
  }
}

// u (S cells per axis) += its parent's value in uc
void prolong_l(const MG& g, int S, float* u, const float* uc) {
  const int C = S * S * S, Sc = S / 2, Cc = Sc * Sc * Sc;
#pragma omp parallel for schedule(static)
  for (int64_t l = 0; l < g.L; ++l)
    for (int c = 0; c < C; ++c) {
      const int x = c / (S * S), y = (c / S) % S, z = c % S;
      u[l * C + c] += uc[l * Cc + ((x >> 1) * Sc + (y >> 1)) * Sc + (z >> 1)];
    }
}

// level 3 (one cell per leaf) colour sweep
void sweep3(const MG& g, const float* b, float* u, float inv_s, int colour, bool first, float omega) {
#pragma omp parallel for schedule(static)
  for (int64_t l = 0; l < g.L; ++l) {
    if (g.lcol[l] != colour) {
      if (first && colour == 0) u[l] = 0.f;
      continue;
    }
    const float s = (first && colour == 0) ? 0.f : nsum(g, 1, u, l, 0, 0, 0);
    const float gs = (b[l] * inv_s + s) * (1.0f / 6.0f);
    u[l] = first ? omega * gs : std::fma(omega, gs - u[l], u[l]);
  }
}

// the coarsest levels: level 3 with scale 64 (right-hand side b3, solution u3)
void coarse(MG& g) {
  const Shape& M = g.M;
  const float inv = 1.0f / 64.0f;
  float* b3 = g.b[3].data();
  float* u3 = g.u[3].data();
  if (!g.two) {
    for (int k = 0; k < M.coarse; ++k) {
      sweep3(g, b3, u3, inv, 0, k == 0, M.coarse_omega);
      sweep3(g, b3, u3, inv, 1, k == 0, M.coarse_omega);
    }
    return;
  }
  for (int k = 0; k < M.c3_pre; ++k) {
    sweep3(g, b3, u3, inv, 0, k == 0, M.c3_omega);
    sweep3(g, b3, u3, inv, 1, k == 0, M.c3_omega);
  }
  for (int64_t l = 0; l < g.L; ++l) g.rho3[l] = std::fma(-6.0f, u3[l], b3[l] * inv + nsum(g, 1, u3, l, 0, 0, 0));
  for (int64_t k = 0; k < g.B; ++k) {
    float acc = 0.f;
    for (int q = 0; q < 8; ++q) acc += g.child[k][q] < 0 ? 0.f : g.rho3[g.child[k][q]];
    g.b4[k] = M.c4_over * acc;
  }
  // level 4: red-black SOR from zero
  std::vector<float>& e = g.e4;
  for (int it = 0; it < M.c4_iters; ++it)
    for (int colour = 0; colour < 2; ++colour)
      for (int64_t k = 0; k < g.B; ++k) {
        if (g.bcol[k] != colour) continue;
        float acc = g.b4[k];
        if (!(it == 0 && colour == 0))
          for (int q = 0; q < 6; ++q) {
            const int32_t nb = g.bnb[k][q];
            acc = std::fma((float)g.cnt[k][q], nb < 0 ? 0.f : e[nb], acc);
          }
        const float gs = acc * g.invd[k];
        e[k] = it == 0 ? M.c4_omega * gs : std::fma(M.c4_omega, gs - e[k], e[k]);
      }
  for (int64_t l = 0; l < g.L; ++l) u3[l] = u3[l] + e[g.blk[l]];
  for (int k = 0; k < M.c3_post; ++k) {
    sweep3(g, b3, u3, inv, 1, false, M.c3_omega_post);
    sweep3(g, b3, u3, inv, 0, false, M.c3_omega_post);
  }
}

// z = V(r) on leaf-dense level-0 vectors
void vcycle(MG& g, const float* r, float* z) {
  const Shape& M = g.M;
  const float c = M.jacobi0 * 0.166666667f;  // pointwise pre-smoothing from zero: z' = c r
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < g.L * 512; ++i) z[i] = c * r[i];
  restrict_l(g, 8, r, z, 1.0f, M.over, g.b[1].data());
  for (int l = 1; l <= 2; ++l) {
    const int S = 8 >> l;
    const float scale = (float)(1 << (2 * l));
    sweep_l(g, S, g.b[l].data(), g.u[l].data(), 1.0f / scale, 0, true, M.omega_l);
    sweep_l(g, S, g.b[l].data(), g.u[l].data(), 1.0f / scale, 1, true, M.omega_l);
    restrict_l(g, S, g.b[l].data(), g.u[l].data(), scale, l == 2 ? M.over3 : M.over_deep, g.b[l + 1].data());
  }
  coarse(g);
  for (int l = 2; l >= 1; --l) {
    const int S = 8 >> l;
    const float scale = (float)(1 << (2 * l));
    prolong_l(g, S, g.u[l].data(), g.u[l + 1].data());
    for (int k = 0; k < (l == 2 ? M.post2 : 1); ++k) {
      sweep_l(g, S, g.b[l].data(), g.u[l].data(), 1.0f / scale, 1, false, M.omega_l);
      sweep_l(g, S, g.b[l].data(), g.u[l].data(), 1.0f / scale, 0, false, M.omega_l);
    }
  }
  prolong_l(g, 8, z, g.u[1].data());
  sweep0(g, r, z, 1, false, M.omega0);
  sweep0(g, r, z, 0, false, M.omega0);
}

MG* create(const int32_t* origins, int64_t L) {
  auto* g = new MG();
  g->L = L;
  g->lc.resize(L);
  g->fn.resize(L);
  g->lcol.resize(L);
  auto key = [](int64_t x, int64_t y, int64_t z) {
    return ((uint64_t)(x + (1 << 20)) << 42) | ((uint64_t)(y + (1 << 20)) << 21) | (uint64_t)(z + (1 << 20));
  };
  std::unordered_map<uint64_t, int32_t> at;
  for (int64_t l = 0; l < L; ++l) {
    for (int a = 0; a < 3; ++a) g->lc[l][a] = origins[3 * l + a] >> 3;
    at[key(g->lc[l][0], g->lc[l][1], g->lc[l][2])] = (int32_t)l;
    g->lcol[l] = (((origins[3 * l] + origins[3 * l + 1] + origins[3 * l + 2]) >> 3) & 1);
  }
  for (int64_t l = 0; l < L; ++l)
    for (int q = 0; q < 6; ++q) {
      int64_t c[3] = {g->lc[l][0], g->lc[l][1], g->lc[l][2]};
      c[q >> 1] += (q & 1) ? -1 : 1;
      auto it = at.find(key(c[0], c[1], c[2]));
      g->fn[l][q] = it == at.end() ? -1 : it->second;
    }
  // level 4 blocks
  std::unordered_map<uint64_t, int32_t> bid;
  g->blk.resize(L);
  for (int64_t l = 0; l < L; ++l) {
    const int32_t bx = g->lc[l][0] >> 1, by = g->lc[l][1] >> 1, bz = g->lc[l][2] >> 1;
    auto it = bid.emplace(key(bx, by, bz), (int32_t)g->child.size());
    if (it.second) {
      g->child.emplace_back();
      g->child.back().fill(-1);
      g->bcol.push_back((bx + by + bz) & 1);
    }
    g->blk[l] = it.first->second;
    g->child[g->blk[l]][((g->lc[l][0] & 1) << 2) | ((g->lc[l][1] & 1) << 1) | (g->lc[l][2] & 1)] = (int32_t)l;
  }
  g->B = (int64_t)g->child.size();
  g->bnb.assign(g->B, {-1, -1, -1, -1, -1, -1});
  g->cnt.assign(g->B, {0, 0, 0, 0, 0, 0});
  std::vector<int>& D = g->D;
  D.assign(g->B, 0);
  for (int64_t l = 0; l < L; ++l) {
    const int32_t b = g->blk[l];
    int inside = 0;
    for (int q = 0; q < 6; ++q) {
      const int32_t nl = g->fn[l][q];
      if (nl < 0) continue;
      if (g->blk[nl] == b) {
        ++inside;
      } else {
        ++g->cnt[b][q];
        g->bnb[b][q] = g->blk[nl];
      }
    }
    D[b] += 6 - inside;
  }
  g->invd.resize(g->B);
  for (int64_t k = 0; k < g->B; ++k) g->invd[k] = 1.0f / (float)D[k];
  int64_t nc[2] = {0, 0};
  for (int64_t k = 0; k < g->B; ++k) ++nc[g->bcol[k]];
  g->two = std::max(nc[0], nc[1]) <= g->M.max_blocks_per_colour;
  for (int l = 1; l <= 3; ++l) {
    const int64_t n = L << (3 * (3 - l));
    g->b[l].assign(n, 0.f);
    g->u[l].assign(n, 0.f);
  }
  g->rho3.assign(L, 0.f);
  g->b4.assign(g->B, 0.f);
  g->e4.assign(g->B, 0.f);
  return g;
}

void destroy(MG* g) { delete g; }

// A d on level 0 (6 d - the face neighbours, zero outside): the library's product order
void apply_A(const MG& g, const float* d, float* q) {
#pragma omp parallel for schedule(static)
  for (int64_t l = 0; l < g.L; ++l)
    for (int v = 0; v < 512; ++v) {
      const int64_t i = l * 512 + v;
      q[i] = 6.0f * d[i] - nsum(g, 8, d, l, v >> 6, (v >> 3) & 7, v & 7);
    }
}
double dot(int64_t n, const float* a, const float* b) {
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (int64_t i = 0; i < n; ++i) s += (double)a[i] * (double)b[i];
  return s;
}

// Flexible CG with truncation 1 on the V-cycle, for A p = -div dx^2 from p = 0 (the library's
// HNS_SOLVER_MGCG): d = z + beta d_prev with beta = -(z . q_prev) / (d_prev . q_prev), q = A d,
// alpha = (z . r) / (d . q); r replaced by b - A x (FP64-evaluated) every 64 iterations and before
// a stop is accepted. Returns the iterations; rel: ||b - A p|| / ||b|| at the end.
int mgcg(MG& g, const float* div, float dx, double tol, int max_iters, float* p, double* rel) {
  const int64_t n = g.L * 512;
  const float dx2 = dx * dx;
  std::vector<float> r(n), z(n), d(n), q(n), bvec(n);
  for (int64_t i = 0; i < n; ++i) {
    bvec[i] = -div[i] * dx2;
    r[i] = bvec[i];
    p[i] = 0.f;
  }
  const double bb = dot(n, bvec.data(), bvec.data());
  auto replace = [&]() {
#pragma omp parallel for schedule(static)
    for (int64_t l = 0; l < g.L; ++l)
      for (int v = 0; v < 512; ++v) {
        const int64_t i = l * 512 + v;
        double s = 0.0;
        for (int qf = 0; qf < 6; ++qf) {
          const int64_t j = face_cell(g, 8, l, v >> 6, (v >> 3) & 7, v & 7, qf);
          s += j < 0 ? 0.0 : (double)p[j];
        }
        r[i] = (float)((double)bvec[i] - (6.0 * (double)p[i] - s));
      }
  };
  const double tol2 = tol * tol;
  double sigma_prev = 0.0;
  int it = 0, replaced_at = 0;
  while (it < max_iters) {
    if (it > 0 && it % 64 == 0 && replaced_at != it) {
      replace();
      replaced_at = it;
    }
    const double rr = dot(n, r.data(), r.data());
    if (!(rr > tol2 * bb)) {
      if (it == replaced_at) break;
      replace();  // the stop stands only on the replaced residual
      replaced_at = it;
      continue;
    }
    std::fill(z.begin(), z.end(), 0.f);
    vcycle(g, r.data(), z.data());
    if (it == 0) {
      std::copy(z.begin(), z.end(), d.begin());
    } else {
      const float beta = (float)(-dot(n, z.data(), q.data()) / sigma_prev);
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < n; ++i) d[i] = std::fma(beta, d[i], z[i]);
    }
    const double zr = dot(n, z.data(), r.data());
    apply_A(g, d.data(), q.data());
    const double sigma = dot(n, d.data(), q.data());
    if (!(sigma > 0.0)) break;
    const float alpha = (float)(zr / sigma);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      p[i] = std::fma(alpha, d[i], p[i]);
      r[i] = std::fma(-alpha, q[i], r[i]);
    }
    sigma_prev = sigma;
    ++it;
  }
  replace();
  const double rr = dot(n, r.data(), r.data());
  *rel = bb > 0.0 ? std::sqrt(rr / bb) : 0.0;
  return it;
}

}  // namespace oracle_mg

extern "C" {

void* oracle_mg_create(const int32_t* origins, int64_t L) { return oracle_mg::create(origins, L); }
void oracle_mg_destroy(void* g) { delete static_cast<oracle_mg::MG*>(g); }
int oracle_mg_two_level(void* g) { return static_cast<oracle_mg::MG*>(g)->two ? 1 : 0; }
int64_t oracle_mg_blocks(void* g) { return static_cast<oracle_mg::MG*>(g)->B; }
// z = V(r); r and z leaf-dense (512 L floats)
void oracle_mg_vcycle(void* g, const float* r, float* z) {
  auto* m = static_cast<oracle_mg::MG*>(g);
  std::fill(z, z + m->L * 512, 0.f);
  oracle_mg::vcycle(*m, r, z);
}
// the level operators for the Galerkin checks: out = scale (6 u - face neighbours) at level l
// (0..3; S = 8 >> l), and level 4's operator D_B e - sum n_BB' e_B'
void oracle_mg_apply_level(void* g, int level, const float* u, float scale, float* out) {
  auto* m = static_cast<oracle_mg::MG*>(g);
  const int S = 8 >> level, C = S * S * S;
  for (int64_t l = 0; l < m->L; ++l)
    for (int c = 0; c < C; ++c) {
      const int x = c / (S * S), y = (c / S) % S, z = c % S;
      const int64_t i = l * C + c;
      out[i] = scale * (6.0f * u[i] - oracle_mg::nsum(*m, S, u, l, x, y, z));
    }
}
void oracle_mg_apply_level4(void* g, const double* e, double* out) {
  auto* m = static_cast<oracle_mg::MG*>(g);
  for (int64_t k = 0; k < m->B; ++k) {
    double s = (double)m->D[k] * e[k];
    for (int q = 0; q < 6; ++q)
      if (m->bnb[k][q] >= 0) s -= (double)m->cnt[k][q] * e[m->bnb[k][q]];
    out[k] = s;
  }
}
// level-4 block of every leaf (for building P), and each block's colour
void oracle_mg_blocks_of(void* g, int32_t* blk, int32_t* colour) {
  auto* m = static_cast<oracle_mg::MG*>(g);
  for (int64_t l = 0; l < m->L; ++l) blk[l] = m->blk[l];
  for (int64_t k = 0; k < m->B; ++k) colour[k] = m->bcol[k];
}
int oracle_mgcg(void* g, const float* div, float dx, double tol, int max_iters, float* p, double* rel) {
  return oracle_mg::mgcg(*static_cast<oracle_mg::MG*>(g), div, dx, tol, max_iters, p, rel);
}

}  // extern "C"
