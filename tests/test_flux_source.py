"""Run-time compiled flux functors (no GPU needed): the library's kernels
compile with the example functors for every order, both kernels and both
bases, and a source that does not compile is refused with the compiler's
message."""
import pytest

from wavedg_amd import capi, fluxes


@pytest.mark.parametrize("n", [2, 3, 4, 5, 6, 7, 8, 9])
@pytest.mark.parametrize("which", [0, 1])
def test_euler_source_compiles(n, which):
    capi.flux_check(fluxes.EULER_LLF, n, 4, capi.GAUSS_LOBATTO, which)


@pytest.mark.parametrize("n", [2, 9])
@pytest.mark.parametrize("which", [0, 1])
def test_sources_compile_on_legendre_basis(n, which):
    capi.flux_check(fluxes.EULER_LLF, n, 4, capi.GAUSS_LEGENDRE, which)
    capi.flux_check(fluxes.WAVE_UPWIND, n, 3, capi.GAUSS_LEGENDRE, which)


def test_only_the_used_functor_is_needed():
    src = "__device__ void flux(const double* x, const double* u, double* F) { F[0] = u[0]; F[1] = 0.0; }"
    capi.flux_check(src, 4, 1, capi.GAUSS_LOBATTO, 0)
    with pytest.raises(capi.WdgError, match="numerical_flux"):
        capi.flux_check(src, 4, 1, capi.GAUSS_LOBATTO, 1)


def test_bad_source_reports_compiler_log():
    with pytest.raises(capi.WdgError, match="undeclared identifier 'not_defined'"):
        capi.flux_check("__device__ void flux(const double* x, const double* u, double* F) { F[0] = not_defined; }",
                        3, 1, capi.GAUSS_LOBATTO, 0)
    with pytest.raises(capi.WdgError):
        capi.flux_check(fluxes.EULER_LLF, 10, 4, capi.GAUSS_LOBATTO, 0)


def test_largest_collocated_case_compiles():
    """n = 9, 8 variables: the collocated kernel compiles; the quadrature
    form's 2n-point rule does not fit one workgroup's local memory, so the
    check skips it (and Flux(..., approx_quad=False) refuses it)."""
    src = ("__device__ void flux(const double* x, const double* u, double* F)"
           " { for (int d = 0; d < 8; ++d) { F[d] = u[d]; F[8 + d] = 0.5 * u[d]; } }")
    capi.flux_check(src, 9, 8, capi.GAUSS_LOBATTO, 0)
