// pn_chunkmap.cpp -- chunk id <-> chunk coordinate map.
//
// The reference numbers chunks along a generalized Hilbert curve ("gilbert",
// J. Cerveny; nix/sfc.cpp:31-141,200-509) so that, with seed_type = 'fixed',
// each chunk's random seed is its id.  Reproducing the same numbering keeps the
// seeded initial conditions identical to the reference's.
#include <cstdlib>
#include <vector>

namespace picnix {

namespace {

int sgn(int v) { return (v > 0) - (v < 0); }

struct Curve {
  int               nz, ny, nx;
  std::vector<int>* out;
  int               next = 0;

  void emit(int x, int y, int z) { (*out)[(z * ny + y) * nx + x] = next++; }

  // Planar curve over the rectangle spanned from (x, y) by the axes a and b
  // (nix/sfc.cpp:200-298).  Unlike the box walk, each branch nudges only the
  // half it splits.
  void walk2(int x, int y, const int a[2], const int b[2])
  {
    int w = std::abs(a[0] + a[1]), h = std::abs(b[0] + b[1]);
    int da[2] = {sgn(a[0]), sgn(a[1])}, db[2] = {sgn(b[0]), sgn(b[1])};
    if (h == 1 || w == 1) {
      const int* dv  = h == 1 ? da : db;
      int        len = h == 1 ? w : h;
      for (int i = 0; i < len; i++) {
        emit(x, y, 0);
        x += dv[0];
        y += dv[1];
      }
      return;
    }
    int a2[2] = {a[0] / 2, a[1] / 2}, b2[2] = {b[0] / 2, b[1] / 2};
    if (2 * w > 3 * h) {
      if ((std::abs(a2[0] + a2[1]) % 2) && w > 2) {
        a2[0] += da[0];
        a2[1] += da[1];
      }
      int a3[2] = {a[0] - a2[0], a[1] - a2[1]};
      walk2(x, y, a2, b);
      walk2(x + a2[0], y + a2[1], a3, b);
    } else {
      if ((std::abs(b2[0] + b2[1]) % 2) && h > 2) {
        b2[0] += db[0];
        b2[1] += db[1];
      }
      int a3[2]  = {a[0] - a2[0], a[1] - a2[1]};
      int b3[2]  = {b[0] - b2[0], b[1] - b2[1]};
      int nb2[2] = {-b2[0], -b2[1]}, na3[2] = {-a3[0], -a3[1]};
      walk2(x, y, b2, a2);
      walk2(x + b2[0], y + b2[1], a, b3);
      walk2(x + b2[0] + a[0] - da[0] - db[0], y + b2[1] + a[1] - da[1] - db[1], nb2, na3);
    }
  }

  // Walk the box spanned from (x,y,z) by the major axis a, and the minor axes b, c.
  void walk(int x, int y, int z, int a[3], int b[3], int c[3])
  {
    int w = std::abs(a[0] + a[1] + a[2]);
    int h = std::abs(b[0] + b[1] + b[2]);
    int d = std::abs(c[0] + c[1] + c[2]);
    int da[3] = {sgn(a[0]), sgn(a[1]), sgn(a[2])};
    int db[3] = {sgn(b[0]), sgn(b[1]), sgn(b[2])};
    int dc[3] = {sgn(c[0]), sgn(c[1]), sgn(c[2])};

    // straight runs
    int* run = nullptr;
    int  len = 0;
    if (h == 1 && d == 1) {
      run = da;
      len = w;
    } else if (d == 1 && w == 1) {
      run = db;
      len = h;
    } else if (w == 1 && h == 1) {
      run = dc;
      len = d;
    }
    if (run) {
      for (int i = 0; i < len; i++) {
        emit(x, y, z);
        x += run[0];
        y += run[1];
        z += run[2];
      }
      return;
    }

    // halves, nudged so that an odd split of a long side stays odd
    int a2[3], b2[3], c2[3];
    for (int k = 0; k < 3; k++) {
      a2[k] = a[k] / 2;
      b2[k] = b[k] / 2;
      c2[k] = c[k] / 2;
    }
    auto nudge = [](int v2[3], const int dv[3], int full) {
      int half = std::abs(v2[0] + v2[1] + v2[2]);
      if ((half % 2) && full > 2)
        for (int k = 0; k < 3; k++)
          v2[k] += dv[k];
    };
    nudge(a2, da, w);
    nudge(b2, db, h);
    nudge(c2, dc, d);
    int a3[3], b3[3], c3[3];
    for (int k = 0; k < 3; k++) {
      a3[k] = a[k] - a2[k];
      b3[k] = b[k] - b2[k];
      c3[k] = c[k] - c2[k];
    }
    auto neg = [](const int v[3], int r[3]) {
      for (int k = 0; k < 3; k++)
        r[k] = -v[k];
    };
    int p[3] = {x, y, z};
    auto mv  = [&](const int v[3], int sign) {
      for (int k = 0; k < 3; k++)
        p[k] += sign * v[k];
    };

    if (2 * w > 3 * h && 2 * w > 3 * d) {
      walk(p[0], p[1], p[2], a2, b, c);
      mv(a2, +1);
      walk(p[0], p[1], p[2], a3, b, c);
    } else if (3 * h > 4 * d) {
      int nb2[3], na3[3];
      neg(b2, nb2);
      neg(a3, na3);
      walk(p[0], p[1], p[2], b2, c, a2);
      mv(b2, +1);
      walk(p[0], p[1], p[2], a, b3, c);
      mv(a, +1);
      mv(da, -1);
      mv(db, -1);
      walk(p[0], p[1], p[2], nb2, c, na3);
    } else if (3 * d > 4 * h) {
      int nc2[3], na3[3];
      neg(c2, nc2);
      neg(a3, na3);
      walk(p[0], p[1], p[2], c2, a2, b);
      mv(c2, +1);
      walk(p[0], p[1], p[2], a, b, c3);
      mv(a, +1);
      mv(da, -1);
      mv(dc, -1);
      walk(p[0], p[1], p[2], nc2, na3, b);
    } else {
      int nb2[3], nc3[3], nc[3], na3[3];
      neg(b2, nb2);
      neg(c3, nc3);
      neg(c, nc);
      neg(a3, na3);
      walk(p[0], p[1], p[2], b2, c2, a2);
      mv(b2, +1);
      walk(p[0], p[1], p[2], c, a2, b3);
      mv(c, +1);
      mv(db, -1);
      mv(dc, -1);
      walk(p[0], p[1], p[2], a, nb2, nc3);
      mv(a, +1);
      mv(da, -1);
      mv(db, +1);
      walk(p[0], p[1], p[2], nc, na3, b3);
      mv(c, -1);
      mv(db, -1);
      mv(dc, +1);
      walk(p[0], p[1], p[2], nb2, c2, na3);
    }
  }
};

} // namespace

// id_of_coord[(cz * Cy + cy) * Cx + cx]; layout of nix/sfc.cpp:100-141 (get_map3d)
std::vector<int> make_chunk_map(int Cz, int Cy, int Cx)
{
  std::vector<int> id(Cz * Cy * Cx, 0);
  int              ndim = (Cz > 1) + (Cy > 1) + (Cx > 1);
  if (ndim <= 1) {
    for (int i = 0; i < Cz * Cy * Cx; i++)
      id[i] = i;
    return id;
  }
  if (ndim == 3) {
    Curve k{Cz, Cy, Cx, &id};
    int   X[3] = {Cx, 0, 0}, Y[3] = {0, Cy, 0}, Z[3] = {0, 0, Cz};
    if (Cx >= Cy && Cx >= Cz)
      k.walk(0, 0, 0, X, Y, Z);
    else if (Cy >= Cx && Cy >= Cz)
      k.walk(0, 0, 0, Y, X, Z);
    else
      k.walk(0, 0, 0, Z, X, Y);
    return id;
  }
  // two chunked dimensions: a planar curve (nix/sfc.cpp:43-55) on the (fast, slow) pair
  int nf, ns; // extents of the faster and slower of the two chunked axes
  if (Cz == 1) {
    nf = Cx;
    ns = Cy;
  } else if (Cy == 1) {
    nf = Cx;
    ns = Cz;
  } else {
    nf = Cy;
    ns = Cz;
  }
  std::vector<int> plane(nf * ns, 0);
  Curve            k{1, ns, nf, &plane};
  int              F[2] = {nf, 0}, S[2] = {0, ns};
  if (nf >= ns)
    k.walk2(0, 0, F, S);
  else
    k.walk2(0, 0, S, F);
  // plane[(slow * nf) + fast] -> id[(cz * Cy + cy) * Cx + cx]
  for (int s = 0; s < ns; s++)
    for (int f = 0; f < nf; f++)
      id[s * nf + f] = plane[s * nf + f];
  return id;
}

} // namespace picnix
