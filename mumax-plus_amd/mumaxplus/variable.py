"""Variable (mumaxplus/variable.py of the reference)."""

import numpy as _np

from .fieldquantity import FieldQuantity


class Variable(FieldQuantity):
    """Represent a physical variable field, e.g. magnetization."""

    def set(self, value):
        if callable(value):
            X, Y, Z = self.meshgrid
            value = _np.vectorize(value, otypes=[float] * self.shape[0])(X, Y, Z)
        self._impl.set(value)

    def get(self):
        return self._impl.get()
