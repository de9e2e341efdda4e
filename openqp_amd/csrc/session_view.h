// session_view.h -- read-only view of a session's basis and geometry for the
// translation units outside session.hip (the XC build), and the hook that
// frees their per-session state with the session.
#pragma once
#include <hip/hip_runtime.h>

struct oqph_session;

namespace oqph {

struct SessionView {
  int device;
  hipStream_t stream;
  int nsh, nbf, natm;
  const int *am, *harm, *atom, *nprim, *poff, *aooff;
  const double *ex, *cc;   // exponents and contraction coefficients (primitive-normalised)
  const double *xyz, *zn;  // natm x 3 bohr, nuclear charges
};

// false when the session has no basis or geometry yet
bool session_view(oqph_session* s, SessionView* v);
// frees the XC state of s (called by oqph_session_destroy)
void xc_release(oqph_session* s);

}  // namespace oqph
