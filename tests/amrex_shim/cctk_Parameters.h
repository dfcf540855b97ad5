// Test shim: DECLARE_CCTK_PARAMETERS for the three CarpetX parameters the
// Interpolater adapter reads (CarpetX/param.ccl: prolongation_type,
// prolongation_order, prolongate_per_group), bound to globals the check
// program sets.
#pragma once

namespace cctk_shim {
inline const char *&prolongation_type() {
  static const char *v = "natural";
  return v;
}
inline int &prolongation_order() {
  static int v = 1;
  return v;
}
inline int &prolongate_per_group() {
  static int v = 0;
  return v;
}
} // namespace cctk_shim

#define DECLARE_CCTK_PARAMETERS                                                \
  const char *const prolongation_type = cctk_shim::prolongation_type();       \
  const int prolongation_order = cctk_shim::prolongation_order();             \
  const int prolongate_per_group = cctk_shim::prolongate_per_group();         \
  (void)prolongation_type;                                                    \
  (void)prolongation_order;                                                   \
  (void)prolongate_per_group
