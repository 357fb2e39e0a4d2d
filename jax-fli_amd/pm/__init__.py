from .correction import AbstractCorrection, NoCorrection, PGDKernel, SharpeningKernel
from .integrate import count_steps, integrate, time_grid
from .interp import AbstractInterp, DriftInterp, InterpTilerState, NoInterp, OnionTiler, TelephotoInterp
from .lpt import lpt
from .nbody import nbody
from .solvers import AbstractNBodySolver, EfficientDriftDoubleKick, NBodyState, ReversibleDoubleKickDrift, StepOp
