"""StrayField (mumaxplus/strayfield.py of the reference)."""

from . import _mumaxpluscpp as _cpp
from .fieldquantity import FieldQuantity


class StrayField(FieldQuantity):
    """The stray field (T) of a magnet on a grid.

    magnet : the source (Ferromagnet or Antiferromagnet)
    grid   : Grid on which the field is evaluated
    """

    def __init__(self, magnet, grid):
        super().__init__(_cpp.StrayField(magnet._impl, grid._impl))

    @classmethod
    def _from_impl(cls, impl):
        sf = cls.__new__(cls)
        sf._impl = impl
        return sf

    def set_method(self, method):
        """"fft" (default) or "brute" (direct sum)."""
        self._impl.set_method(method)
