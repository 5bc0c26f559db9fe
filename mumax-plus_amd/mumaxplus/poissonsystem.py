"""PoissonSystem (mumaxplus/poissonsystem.py of the reference): the solver of
a ferromagnet's electrical potential from its applied_potential and
conductivity (with AMR)."""


class PoissonSystem:
    """Poisson system which can be solved for the electrical potential."""

    def __init__(self, impl):
        self._impl = impl

    def solve(self):
        """Solve the Poisson equation; returns the potential (1, nz, ny, nx)."""
        return self._impl.solve()

    def set_method(self, method_name):
        """jacobi, conjugategradient, minimalresidual or steepestdescent."""
        return self._impl.solver.set_method(method_name)

    @property
    def max_iter(self):
        """Maximum number of iterations of the solver (no maximum if < 0)."""
        return self._impl.solver.max_iter

    @max_iter.setter
    def max_iter(self, max_iter):
        self._impl.solver.max_iter = max_iter

    @property
    def tol(self):
        """Tolerance on the max norm of the residual."""
        return self._impl.solver.tol

    @tol.setter
    def tol(self, tol):
        self._impl.solver.tol = tol

    def max_norm_residual(self):
        """The max norm of the residual."""
        return self._impl.solver.max_norm_residual()

    @property
    def _solver(self):
        return self._impl.solver

    def _init(self):
        """Build (or rebuild) the linear system."""
        self._impl.init()
