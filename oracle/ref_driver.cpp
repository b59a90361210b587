// ref_driver.cpp -- extern "C" shims that call the REFERENCE's own functions,
// compiled from the reference sources where they lie (see oracle/Makefile,
// target `ref`).  TEST INFRASTRUCTURE ONLY: used to write golden vectors and
// to validate the restatements in mms_oracle.c; never linked into the product.
//
// Reference functions reached from here:
//   divergence_3d / gradient_3d / add_exchange_field   vector_functions.cpp:8-103
//   associated_legendre / spherical_harmonic / factorial numerics.cu:8-114
//   Box::create_tree_bfs / Box::find_neighbors_recurse   Box.cu:200-234
//   LLG_Mprime                                           ode_functions.cpp:11-18
#include "Vector3.hpp"
#include "Cmpx.hpp"
#include "numerics.hpp"
#include "vector_functions.hpp"
#include "Box.hpp"
#include <stdlib.h>

Vector3 LLG_Mprime(const Vector3 &M, const Vector3 &H, const fptype alfa, const fptype gamma, const fptype Ms);

extern "C" {

void ref_divergence_3d(const float *V, int xd, int yd, int zd, float dx, float dy, float dz, float *S) {
    divergence_3d((const Vector3 *)V, xd, yd, zd, dx, dy, dz, S);
}
void ref_gradient_3d(const float *S, int xd, int yd, int zd, float dx, float dy, float dz, float cm, float *V) {
    gradient_3d(S, xd, yd, zd, dx, dy, dz, cm, (Vector3 *)V);
}
void ref_add_exchange_field(const float *M, float Ms, float A, float mu, int xd, int yd, int zd,
                            float dx, float dy, float dz, float *H) {
    add_exchange_field((const Vector3 *)M, Ms, A, mu, xd, yd, zd, dx, dy, dz, (Vector3 *)H);
}
float ref_associated_legendre(int l, int m, float x) { return associated_legendre(l, m, x); }
void ref_spherical_harmonic(int l, int m, float theta, float phi, float *re, float *im) {
    Cmpx c = spherical_harmonic(l, m, theta, phi);
    *re = c.get_re();
    *im = c.get_im();
}
int ref_factorial(int x) { return factorial(x); }
void ref_llg_mprime(const float *M, const float *H, float alfa, float gamma, float Ms, float *out) {
    Vector3 r = LLG_Mprime(Vector3(M[0], M[1], M[2]), Vector3(H[0], H[1], H[2]), alfa, gamma, Ms);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

// Build the reference quadtree to depth `limit` and report, for every box at
// `level` (in the tree's BFS order), its 8 neighbour and 27 interaction-list
// entries as same-level indices (-1 for NULL) and its centre (cx, cy).
int ref_tree_lists(int limit, int level, int *neighbors, int *interactions, float *centres) {
    long total = 0;
    for (int l = 0; l <= limit; l++) total += 1L << (2 * l);
    Box *tree = (Box *)malloc(total * sizeof(Box));
    void **qmem = (void **)malloc((1L << (2 * limit)) * sizeof(void *));
    if (!tree || !qmem) return 1;
    tree[0] = Box(0, 0, limit);
    tree[0].create_tree_bfs(limit, qmem);
    tree[0].find_neighbors_recurse(&tree[0], limit);
    long first = ((1L << (2 * level)) - 1) / 3;  // offset of level's first box
    long nbox = 1L << (2 * level);
    for (long b = 0; b < nbox; b++) {
        Box *n = &tree[first + b];
        for (int j = 0; j < 8; j++)
            neighbors[8 * b + j] = n->neighbor[j] ? (int)((n->neighbor[j] - tree) - first) : -1;
        for (int j = 0; j < 27; j++)
            interactions[27 * b + j] = n->interaction[j] ? (int)((n->interaction[j] - tree) - first) : -1;
        centres[2 * b] = n->cx;
        centres[2 * b + 1] = n->cy;
    }
    free(qmem);
    free(tree);
    return 0;
}

}  // extern "C"
