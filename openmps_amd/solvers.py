"""Named pressure-solver presets of the Computer (used by bench.py and the
validation runs).

Each preset is (mixed precision, warm start, preconditioner of the FP32 CG):
  fp64            the reference algorithm: unpreconditioned FP64 CG from the
                  previous-step pressure (Computer.hpp:1359-1429);
  fp64_zero       the same from a zero guess;
  mixed_jacobi    FP32 CG (Jacobi) inside FP64 iterative refinement;
  mixed_ml        FP32 CG with the multilevel aggregation preconditioner
                  inside FP64 iterative refinement;
  mixed_ml_zero   the same from a zero guess;
  mixed_amg       FP32 CG with the cell-grid multigrid preconditioner (Jacobi
                  plus one V-cycle on the cells' Galerkin operator) inside FP64
                  iterative refinement (mixed_amg_zero: from a zero guess);
  learned_amg     mixed_amg warm-started by the learned model;
  mixed_auto      FP32 CG with the faster of the multilevel and the cell-grid
                  multigrid preconditioner, timed as the flow goes (the first is
                  faster while the flow is young, the second once it has
                  developed) inside FP64 iterative refinement (mixed_auto_zero:
                  from a zero guess; learned_auto: the learned start);
  learned_ml      mixed_ml warm-started by the learned model;
  learned_mixed   mixed_jacobi warm-started by the learned model.
"""
from __future__ import annotations

from . import _lib as L

PRESETS = {
    "fp64": (False, "previous", "none"),
    "fp64_zero": (False, "zero", "none"),
    "fp64_learned": (False, "learned", "none"),
    "mixed": (True, "previous", "none"),
    "mixed_jacobi": (True, "previous", "jacobi"),
    "mixed_ml": (True, "previous", "multilevel"),
    "mixed_ml_zero": (True, "zero", "multilevel"),
    "learned_ml": (True, "learned", "multilevel"),
    "learned_mixed": (True, "learned", "jacobi"),
    "mixed_amg": (True, "previous", "amg"),
    "mixed_amg_zero": (True, "zero", "amg"),
    "learned_amg": (True, "learned", "amg"),
    "mixed_auto": (True, "previous", "auto"),
    "mixed_auto_zero": (True, "zero", "auto"),
    "learned_auto": (True, "learned", "auto"),
}

_WARM = {"previous": L.WARM_PREVIOUS, "zero": L.WARM_ZERO, "learned": L.WARM_LEARNED}
_PREC = {"none": L.PRECOND_NONE, "jacobi": L.PRECOND_JACOBI, "multilevel": L.PRECOND_MULTILEVEL,
         "amg": L.PRECOND_AMG, "auto": L.PRECOND_AUTO}


def configure(computer, preset: str, tol: float, weights: str | None = None, model=None) -> None:
    """Sets the solver of `computer` (stop rule |b - A x| <= tol |b|); a
    learned preset also attaches `model`, or else the model stored at
    `weights` (scenes of one batch share one model)."""
    mixed, warm, prec = PRESETS[preset]
    computer.set_solver(kind=L.SOLVER_MIXED_IR if mixed else L.SOLVER_CG_FP64,
                        stop_rule=L.STOP_RHS_NORM, warm_start=_WARM[warm], precond=_PREC[prec],
                        tol=tol, inner_tol=1e-4)
    if warm == "learned":
        from . import learned
        learned.attach_model(computer, model if model is not None
                             else learned.Model.load(weights or learned.DEFAULT_WEIGHTS))


def is_mixed(preset: str) -> bool:
    return PRESETS[preset][0]


def is_learned(preset: str) -> bool:
    return PRESETS[preset][1] == "learned"


def preconditioner(preset: str) -> str:
    return PRESETS[preset][2]
