"""Module-level ``__getattr__`` for names of the reference package that lie
outside the GPU FTLE path (LSODA, N-D flow maps, LAVD/iLE/Eulerian
diagnostics, S^2 FTLE, hyperbolic/elliptic LCS, jit callables): importing
one raises ImportError naming the name, instead of a bare AttributeError."""


def outside(module):
    def __getattr__(name):
        if name.startswith("__"):
            raise AttributeError(name)
        raise ImportError(
            f"{module}.{name} is outside the FTLE path this package runs on the GPU "
            f"(flow maps, FTLE, Cauchy-Green eigen, FTLE ridges)")

    return __getattr__
