"""Parameter (mumaxplus/parameter.py of the reference)."""

import numpy as _np

from .fieldquantity import FieldQuantity


class Parameter(FieldQuantity):
    """Represent a physical material parameter, e.g. the exchange stiffness."""

    def __repr__(self):
        return super().__repr__().replace("FieldQuantity", "Parameter")

    @property
    def is_uniform(self):
        return self._impl.is_uniform

    @property
    def is_dynamic(self):
        return self._impl.is_dynamic

    @property
    def uniform_value(self):
        return self._impl.uniform_value

    @uniform_value.setter
    def uniform_value(self, value):
        self._impl.set(value)

    def add_time_term(self, term, mask=None):
        self._impl.add_time_term(term, mask)

    def remove_time_terms(self):
        self._impl.remove_time_terms()

    def set(self, value):
        """Set a uniform value (scalar / 3-tuple), a per-cell array or a
        function of the cell position (x, y, z)."""
        if callable(value):
            try:
                value(0)
                is_time_function = True
            except TypeError:
                is_time_function = False
            if is_time_function:
                self.add_time_term(value)
                return
            X, Y, Z = self.meshgrid
            values = _np.vectorize(value, otypes=[float] * self.ncomp)(X, Y, Z)
            if self.ncomp == 1:
                values = [values]
            self._impl.set(_np.asarray(values))
        elif isinstance(value, tuple) and len(value) and callable(value[0]):
            self.add_time_term(*value)
        else:
            self._impl.set(value)

    def set_in_region(self, region_idx, value):
        raise NotImplementedError("regions are not supported by this build")
