// host_math.h -- host-side numerics of libscft_hip.so.
#pragma once
#include <vector>

namespace scft {
bool romberg_weights(int m, double hh, std::vector<double>& c);
bool pairing_weights(int N, std::vector<double>& pw);
bool pairing_weights_into(int N, double* pw);
double exp_profile(double x, double tau);
void target_position(int M, const double* x, double tau, bool mirror, double* phi_t);
void mesh_uniform(int M, double L, double* x);
bool spline_not_a_knot(int n, const double* x, const double* y, std::vector<double>& M);
double spline_eval(int n, const double* x, const double* y, const double* M, double xp);
bool spline_natural(int n, const double* x, const double* y, std::vector<double>& M);
bool wall_values_natural(int M, const double* x, const double* w_interior, double* wall0,
                         double* wallL);
bool wall_values_from_unknowns(int M, const double* x, const double* w_unknowns, bool dd,
                               double* wall0, double* wallL);
std::vector<double> refine_by_gradient(int M, const double* x, const double* w_nodes,
                                       double factor, bool wall_right);
bool free_energy(int M, const double* x, const double* w_nodes, double L, double tau, bool mirror,
                 int k, double f0bar, double* F);
}  // namespace scft
