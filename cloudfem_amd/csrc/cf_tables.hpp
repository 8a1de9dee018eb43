// Reference-element tables of the P1 and Tet10 elements, evaluated at compile
// time and placed in constant memory.
//
// Values follow the reference's definitions (cited per table); the Dunavant
// face rule keeps the reference's 15-digit literals so the face integrals
// round like the reference's.
#pragma once

namespace cf {

// Tet10 local edge (i, j) of midpoint slot 4+s (cloudfem/mesh/tet10_mesh.py:33).
constexpr int kEdge[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};

struct Tet10Tables {
  double w[5];             // Keast weights, sum 1 (tet10.py:84-100)
  double N[5][10];         // shape functions at the points (tet10.py:63-81)
  double D[5][10][3];      // dN/dxi, xi = (L2, L3, L4) (tet10.py:120-123)
  double R[10][10][3][3];  // sum_q w_q D[q][a][m] D[q][b][n]
  double M[10][10];        // int N_a N_b dV / V (tet10.py:185-226)
  double Dc[10][3];        // dN/dxi at the centroid (elastic_tet10.py:255-262)
};

constexpr double fact(int n) {
  double f = 1.0;
  for (int i = 2; i <= n; ++i) f *= i;
  return f;
}

// (1/V) int L^alpha dV = 6 alpha! / (|alpha| + 3)!  (tet10.py:202-208)
constexpr double bary_moment(const int (&a)[4]) {
  return 6.0 * (fact(a[0]) * fact(a[1]) * fact(a[2]) * fact(a[3])) /
         fact(a[0] + a[1] + a[2] + a[3] + 3);
}

constexpr Tet10Tables build_tet10_tables() {
  Tet10Tables t{};
  const double a = 0.5, b = 1.0 / 6.0;
  const double L[5][4] = {{0.25, 0.25, 0.25, 0.25},
                          {a, b, b, b},
                          {b, a, b, b},
                          {b, b, a, b},
                          {b, b, b, a}};
  t.w[0] = -4.0 / 5.0;
  for (int q = 1; q < 5; ++q) t.w[q] = 9.0 / 20.0;
  for (int q = 0; q < 5; ++q) {
    double dL[10][4] = {};
    for (int i = 0; i < 4; ++i) {
      t.N[q][i] = L[q][i] * (2.0 * L[q][i] - 1.0);
      dL[i][i] = 4.0 * L[q][i] - 1.0;
    }
    for (int s = 0; s < 6; ++s) {
      int i = kEdge[s][0], j = kEdge[s][1];
      t.N[q][4 + s] = 4.0 * L[q][i] * L[q][j];
      dL[4 + s][i] = 4.0 * L[q][j];
      dL[4 + s][j] = 4.0 * L[q][i];
    }
    for (int n = 0; n < 10; ++n)
      for (int m = 0; m < 3; ++m) t.D[q][n][m] = dL[n][m + 1] - dL[n][0];
  }
  for (int p = 0; p < 10; ++p)
    for (int r = 0; r < 10; ++r)
      for (int m = 0; m < 3; ++m)
        for (int n = 0; n < 3; ++n) {
          double acc = 0.0;
          for (int q = 0; q < 5; ++q) acc += t.w[q] * t.D[q][p][m] * t.D[q][r][n];
          t.R[p][r][m][n] = acc;
        }
  {
    double dL[10][4] = {};
    for (int i = 0; i < 4; ++i) dL[i][i] = 4.0 * 0.25 - 1.0;
    for (int s = 0; s < 6; ++s) {
      dL[4 + s][kEdge[s][0]] = 4.0 * 0.25;
      dL[4 + s][kEdge[s][1]] = 4.0 * 0.25;
    }
    for (int n = 0; n < 10; ++n)
      for (int m = 0; m < 3; ++m) t.Dc[n][m] = dL[n][m + 1] - dL[n][0];
  }
  // Shape functions as polynomials in L: corner 2 L_i^2 - L_i, edge 4 L_i L_j.
  int ex[10][2][4] = {};
  double co[10][2] = {};
  int nt[10] = {};
  for (int i = 0; i < 4; ++i) {
    ex[i][0][i] = 2;
    co[i][0] = 2.0;
    ex[i][1][i] = 1;
    co[i][1] = -1.0;
    nt[i] = 2;
  }
  for (int s = 0; s < 6; ++s) {
    ex[4 + s][0][kEdge[s][0]] += 1;
    ex[4 + s][0][kEdge[s][1]] += 1;
    co[4 + s][0] = 4.0;
    nt[4 + s] = 1;
  }
  for (int p = 0; p < 10; ++p)
    for (int r = 0; r < 10; ++r) {
      double acc = 0.0;
      for (int u = 0; u < nt[p]; ++u)
        for (int v = 0; v < nt[r]; ++v) {
          int al[4] = {ex[p][u][0] + ex[r][v][0], ex[p][u][1] + ex[r][v][1],
                       ex[p][u][2] + ex[r][v][2], ex[p][u][3] + ex[r][v][3]};
          acc += co[p][u] * co[r][v] * bary_moment(al);
        }
      t.M[p][r] = acc;
    }
  return t;
}

struct FaceTables {
  double w6[6];
  double N6[6][6];  // 6-node triangle functions at the Dunavant points
  double w3[3];
  double N3[3][6];  // ... at the 3-point rule
};

constexpr void face_shape(const double (&l)[3], double (&out)[6]) {
  const int fe[3][2] = {{0, 1}, {1, 2}, {2, 0}};  // tet10.py:293
  for (int i = 0; i < 3; ++i) out[i] = l[i] * (2.0 * l[i] - 1.0);
  for (int s = 0; s < 3; ++s) out[3 + s] = 4.0 * l[fe[s][0]] * l[fe[s][1]];
}

constexpr FaceTables build_face_tables() {
  FaceTables t{};
  // 6-point degree-4 rule, the reference's literals (tet10.py:300-308)
  const double L6[6][3] = {
      {0.445948490915965, 0.445948490915965, 0.108103018168070},
      {0.445948490915965, 0.108103018168070, 0.445948490915965},
      {0.108103018168070, 0.445948490915965, 0.445948490915965},
      {0.091576213509771, 0.091576213509771, 0.816847572980459},
      {0.091576213509771, 0.816847572980459, 0.091576213509771},
      {0.816847572980459, 0.091576213509771, 0.091576213509771}};
  for (int q = 0; q < 6; ++q) {
    t.w6[q] = q < 3 ? 0.223381589678011 : 0.109951743655322;
    const double l[3] = {L6[q][0], L6[q][1], L6[q][2]};
    face_shape(l, t.N6[q]);
  }
  // 3-point degree-2 rule (elastic_tet10.py:50-53)
  for (int q = 0; q < 3; ++q) {
    double l[3] = {1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0};
    l[q] = 2.0 / 3.0;
    t.w3[q] = 1.0 / 3.0;
    face_shape(l, t.N3[q]);
  }
  return t;
}

}  // namespace cf
