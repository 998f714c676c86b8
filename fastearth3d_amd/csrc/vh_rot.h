// vh_rot.h -- device state and host wrappers of the rotation channels.
#pragma once
#include "vh_internal.h"

// scalar slots of rot_state::sc (complex values take two slots: re, im)
enum {
    RS_M = 0, RS_CLOAD = 2,
    RS_DF = 4,        // dF[load], dF[tidal]  (4 slots)
    RS_DU = 8,        // dU[load], dU[tidal]  (4 slots)
    RS_FE_LOAD = 12, RS_UE_TIDAL = 13, RS_KTE = 14, RS_KS = 15,
    RS_A = 16, RS_G = 17, RS_CMA = 18, RS_OMEGA = 19,
    RS_I13 = 20, RS_I23 = 21,
    RS_N = 24
};

struct rot_state {
    double* mem;       // [2][ne][24]
    double* MkPerDt;   // [2][ne]
    double* xU;        // [2][nr]
    double* xV;        // [2][nr]
    double* dN;        // [2][4][nr]
    double* sc;        // [RS_N]
};

struct rot_args {
    int nr, nd;
    const double* r;
    const double* mu;
    // elements that carry channel memory (mu/eta > 0) and their nodes; G2 is
    // the degree-2 solution operator restricted to those nodes' U/V dofs plus
    // the surface U, F, V rows: [2 nmn2 + 3][2 nmn2]
    int nem2, nmn2;
    const int* emap2;    // [nem2]  -> mesh element
    const int* mnode2;   // [nr]    mesh node -> index in mnodes2 or -1
    const int* mnodes2;  // [nmn2]
    const double* G2;
    rot_state rs;
};

void rot_begin(hipStream_t st, const rot_args& a);
void rot_srot(hipStream_t st, const rot_args& a, int nphi, int nlat, const double* sc2, const double* cphi,
              const double* sphi, double* srot);
void rot_inertia(hipStream_t st, const rot_args& a, const vh_sht& s, const double* load, const double* sc2,
                 const double* cphi, const double* sphi, double* part);
void rot_solve_commit(hipStream_t st, const rot_args& a, double dt);

// ---- reverse mode: adjoint state of the rotation channels
enum { RA_M = 0, RA_DF = 2, RA_DU = 6, RA_I = 10, RA_N = 12 };
struct rot_adj {
    double* mem;   // [2][ne][24] adjoint of the channel memory
    double* dN;    // [2][4][nr]  adjoint of the channel drift (cleared by rot_begin_adj)
    double* sc;    // [RA_N]: m, dF (4), dU (4), I13/I23
};
// one sub-interval of solve_m + commit, transposed (m_h in: adjoint of the m it stores)
void rot_commit_adj(hipStream_t st, const rot_args& a, const rot_adj& h, double dt);
// rot_begin transposed: mem_h += (drift)^T (dN_h, dU_h, dF_h), which are cleared
void rot_begin_adj(hipStream_t st, const rot_args& a, const rot_adj& h);
// s_rot transposed, srot_h in density form (adjoint / quadrature weight): m_h, dF_h, dU_h (tidal)
void rot_srot_adj(hipStream_t st, const rot_args& a, const rot_adj& h, const vh_sht& s, const double* srot_h,
                  const double* sc2, const double* cphi, const double* sphi, double* part, double* tot);
// g = gamma (I13_h cos + I23_h sin): the load inertia's grid adjoint divided by the quadrature weight
void rot_inertia_adj(hipStream_t st, const rot_adj& h, int nphi, int nlat, const double* sc2, const double* cphi,
                     const double* sphi, double* g);
