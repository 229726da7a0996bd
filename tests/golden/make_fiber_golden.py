"""Write tests/golden/fiber_golden.npz from the reference's own fiber code.

Run once, in the development container where the reference is mounted:
    python tests/golden/make_fiber_golden.py
It imports python/fiber/fiber.py, python/tstep/fiber_matrices.py and their
python/utils helpers (NumPy/SciPy only) and records, for three seeded curved
fibers, the inputs and the reference outputs of:
  fiber_matrices.compute_matrices / get_matrices  (tstep/fiber_matrices.py:38-153)
  fiber.form_linear_operator                      (fiber/fiber.py:322-447)
  fiber.compute_RHS                               (fiber/fiber.py:473-582)
  fiber.apply_BC_rectangular                      (fiber/fiber.py:810-961),
                                                  free, clamped minus end, hinged plus end
  fiber.force_operator                            (fiber/fiber.py:1399-1431)
Only arrays are stored (no pickles); the reference itself is not copied.
"""
import os
import sys

import numpy as np

REF = "/root/reference/python"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fiber_golden.npz")


def main():
    sys.path.insert(0, REF)
    from fiber import fiber as fiber_mod
    from tstep import fiber_matrices as fm_mod

    data = {}
    cases = [(16, 0.9, 0.02, 5e-3), (24, 1.5, 0.01, 1e-3), (32, 1.0, 0.01, 1e-3)]
    for ci, (N, L, E, dt) in enumerate(cases):
        rng = np.random.default_rng(100 + ci)
        fm = fm_mod.fiber_matrices(num_points=N, num_points_finite_diff=4)
        fm.compute_matrices()
        s = np.linspace(0.0, L, N)
        x = np.stack([0.3 * np.cos(2.0 * s), 0.3 * np.sin(2.0 * s), 0.8 * s], axis=1)
        x += rng.normal(scale=1e-3, size=x.shape)
        fib = fiber_mod.fiber(num_points=N, num_points_finite_diff=4, length=L, epsilon=1e-3,
                              E=E, viscosity=1.3, dt=dt)
        fib.x = x.copy()
        D1, D2, D3, D4 = fm.get_matrices(fib.length_previous, fib.num_points_up, "Ds")
        fib.xs, fib.xss, fib.xsss, fib.xssss = D1 @ x, D2 @ x, D3 @ x, D4 @ x
        force = rng.normal(size=(N, 3))
        flow = rng.normal(size=(N, 3))
        A = fib.form_linear_operator(fm)
        RHS = fib.compute_RHS(fib_mat=fm, force_external=force, flow=flow)
        fib.set_BC(BC_end_vec_0=force[-1])
        A_sq, b = fib.apply_BC_rectangular(A, RHS, fm, flow, force)
        # minus end clamped (a fiber attached to a body, tstep.py:878-881)
        fib.set_BC(BC_start_0="velocity", BC_start_1="angular_velocity",
                   BC_end_vec_0=force[-1])
        A_cl, b_cl = fib.apply_BC_rectangular(A, RHS, fm, flow, force)
        # plus end hinged (a fiber that reached the periphery, tstep.py:886-888)
        fib.set_BC(BC_end_0="velocity", BC_end_1="torque")
        A_hg, b_hg = fib.apply_BC_rectangular(A, RHS, fm, flow, force)
        F = fib.force_operator(fm)
        w, _, _, _ = fm.get_matrices(L, N, "weights_all")
        _, _, P_cheb, _ = fm.get_matrices(L, N, "PX_PT_Pcheb")
        p = f"c{ci}_"
        data.update({p + "N": np.array(N), p + "L": np.array(L), p + "E": np.array(E),
                     p + "dt": np.array(dt), p + "eta": np.array(1.3), p + "eps": np.array(1e-3),
                     p + "x": x, p + "force": force, p + "flow": flow,
                     p + "D1": D1, p + "D2": D2, p + "D3": D3, p + "D4": D4,
                     p + "weights": w, p + "P_cheb": P_cheb, p + "A": A, p + "RHS": RHS,
                     p + "A_sq": A_sq, p + "b": b, p + "F": F,
                     p + "A_sq_clamped": A_cl, p + "b_clamped": b_cl,
                     p + "A_sq_hinged": A_hg, p + "b_hinged": b_hg})
    np.savez_compressed(OUT, **data)
    print(f"wrote {OUT} ({os.path.getsize(OUT) / 1e3:.0f} kB)")


if __name__ == "__main__":
    main()
