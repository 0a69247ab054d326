#!/usr/bin/env bash
# Compiles the reference's own AMIPS energy / Jacobian / Hessian functions
# (src/tetwild/LocalOperations.cpp, the three comformalAMIPS*_new definitions)
# straight from the reference checkout into oracle/_ref/libref_amips.so.
#
# Those three functions use nothing but <cmath> pow, so they compile on their
# own: the recipe cuts their definitions out of LocalOperations.cpp (between the
# Energy definition and LocalOperations::check) and feeds them to g++ on stdin
# after the static member declarations of src/tetwild/LocalOperations.h:109-111.
# Nothing of the reference is written to disk; only the .so lands in _ref/.
# The batch entry points loop over tets on one core (ref_amips_batch_*) or
# over host threads with OpenMP (ref_amips_batch_ejh_mt): the multi-core CPU
# baseline.
# The rest of LocalOperations.cpp needs CGAL, geogram, Eigen and libigl, which
# are not in the image: that part is unbuildable here.
set -euo pipefail
REF=${TETWILD_REFERENCE:-/root/reference}
SRC="$REF/src/tetwild/LocalOperations.cpp"
HERE="$(cd "$(dirname "$0")" && pwd)"
OUT="$HERE/_ref"
[ -f "$SRC" ] || { echo "reference source not found: $SRC" >&2; exit 2; }
mkdir -p "$OUT"
{
  echo '#include <cmath>'
  echo '#include <omp.h>'
  echo 'using std::pow;'
  echo 'namespace tetwild { struct LocalOperations {'
  echo '  static double comformalAMIPSEnergy_new(const double* T);'
  echo '  static void comformalAMIPSJacobian_new(const double* T, double* result_0);'
  echo '  static void comformalAMIPSHessian_new(const double* T, double* result_0); };'
  awk '/^double LocalOperations::comformalAMIPSEnergy_new/{on=1} /^void LocalOperations::check\(\)/{on=0} on' "$SRC"
  echo '}'
  echo 'extern "C" {'
  echo 'double ref_amips_energy(const double* T) { return tetwild::LocalOperations::comformalAMIPSEnergy_new(T); }'
  echo 'void ref_amips_jacobian(const double* T, double* g) { tetwild::LocalOperations::comformalAMIPSJacobian_new(T, g); }'
  echo 'void ref_amips_hessian(const double* T, double* h) { tetwild::LocalOperations::comformalAMIPSHessian_new(T, h); }'
  echo 'void ref_amips_batch_ejh(const double* T, long n, double* E, double* G, double* H) {'
  echo '  for (long i = 0; i < n; ++i) { E[i] = tetwild::LocalOperations::comformalAMIPSEnergy_new(T + 12*i);'
  echo '    tetwild::LocalOperations::comformalAMIPSJacobian_new(T + 12*i, G + 3*i);'
  echo '    tetwild::LocalOperations::comformalAMIPSHessian_new(T + 12*i, H + 9*i); } }'
  echo 'void ref_amips_batch_e(const double* T, long n, double* E) {'
  echo '  for (long i = 0; i < n; ++i) E[i] = tetwild::LocalOperations::comformalAMIPSEnergy_new(T + 12*i); }'
  echo 'void ref_amips_batch_ejh_mt(const double* T, long n, double* E, double* G, double* H, int nt) {'
  echo '  if (nt <= 0) nt = omp_get_max_threads();'
  echo '  _Pragma("omp parallel for schedule(static, 4096) num_threads(nt)")'
  echo '  for (long i = 0; i < n; ++i) { E[i] = tetwild::LocalOperations::comformalAMIPSEnergy_new(T + 12*i);'
  echo '    tetwild::LocalOperations::comformalAMIPSJacobian_new(T + 12*i, G + 3*i);'
  echo '    tetwild::LocalOperations::comformalAMIPSHessian_new(T + 12*i, H + 9*i); } }'
  echo '}'
} | g++ -x c++ -std=c++11 -O3 -DNDEBUG -ffp-contract=off -fopenmp -fPIC -shared -o "$OUT/libref_amips.so" -
echo "built $OUT/libref_amips.so"
