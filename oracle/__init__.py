"""CPU oracle for the f(A)b engine -- TEST INFRASTRUCTURE ONLY.

This package is a plain NumPy/mpmath restatement of the reference's MATLAB
CPU path (``nicksmethodcpu.m``, ``lanczos.m``, ``jacobi_poly_eval.m``,
``polynomial_coeffs.m``, the step loops of ``problem_startup_CPU*.m``) and of
the reference's GPU-side semantics (``lanczos_impl.cu``, ``update_U.cu``,
``update_U_monodomain.cu``).  It exists to *check* the HIP product path.

Rules:

* Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
  ``bench.py`` may import it, and only as the checker or as the timed CPU
  baseline.  The product package ``gpumatfun_amd`` never imports it and never
  falls back to it.
* Every function cites the reference file:line it restates.

Pinning status (what independent check pins each function; see the tests named):

======================  ======================================================
function                pinned by
======================  ======================================================
polynomial_coeffs       exact rationals from the least-squares residual
                        definition; degree 0 -> [1]; s(0)=... closed form;
                        ``tests/test_oracle.py::test_poly_*``
contour_nodes           scalar identity f(lam) = Im sum_j w_j/(xi_j - lam) on
                        [l1, lN] to 1e-12 (mpmath at 40 digits);
                        ``test_contour_scalar_identity``
shifted_gammas          identity q_z(lam)(z-lam) = z p(z) - lam p(lam);
                        ``test_shifted_gamma_identity``
lanczos                 m = N reproduces dense eigh(A p(A)); V^T V = I and
                        V^T A p(A) V = T for small m; ``test_lanczos_*``
tridiag_solve_pivoted   dense numpy.linalg.solve on random complex systems;
                        ``test_tridiag_*``
jacobi_poly_eval        dense polynomial in A on random complex inputs;
                        ``test_poly_eval_dense``
nicksmethodcpu          dense f(A) b (eigh) on N <= 64 and the DST-I closed form
                        for the 2D Dirichlet Laplacian; ``test_fab_*``
step loops              dense exponential-Euler recurrence on tiny N (Fisher,
                        FitzHugh-Nagumo, Mitchell-Schaeffer; each with planted
                        slips -- recovery variable or gate advanced with the
                        old u, current with the old gate -- shown to fail);
                        ``test_step_dense_recurrence``,
                        ``test_fhn_step_loop_dense_recurrence``,
                        ``test_ms_step_loop_dense_recurrence``
ms_reaction             forward-Euler gate in closed form for constant u
                        (h_n = 1 - (1-h0)(1-dt/tau_open)^n), resting state a
                        fixed point, J sign/zeros at u = 0, 1;
                        ``test_ms_reaction_*``
adjoint.gradients       central differences in kappa, b and alpha;
                        ``tests/test_adjoint_oracle.py``
gpu_semantics pipeline  equal to the CPU step loop when the clamp is off and
                        u >= 0; its own clamp/deflation algebra checked densely;
                        update_U_monodomain in its GPU parametrisation against
                        the dense FitzHugh-Nagumo recurrence;
                        ``test_gpu_semantics_*``
C restatement           the NumPy oracle (bitwise-independent of thread count);
                        its long-double build and arithmetic variants are
                        measurement tools, compared to the same
                        ``test_c_restatement_*``
======================  ======================================================
"""

from .poly import polynomial_coeffs, polynomial_coeffs_exact  # noqa: F401
from .contour import contour_setup, shifted_gammas, pack_shifts  # noqa: F401
from .krylov import (  # noqa: F401
    lanczos, jacobi_poly_eval, tridiag_solve_pivoted, nicksmethodcpu,
)
from .stepping import (  # noqa: F401
    cpu_steps, cpu_steps_nonuniform, cpu_steps_monodomain, cpu_steps_ms, ms_reaction,
    GpuSemantics,
)
