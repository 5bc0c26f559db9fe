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
        """Add a time-dependent term (parameter.py:61-113 of the reference).

        The value becomes static + term(t), or static + term(t) * mask per
        cell when a mask is given; terms add up.  term: (t) -> float, or
        (t) -> 3 floats for a vector parameter.  mask: an array of shape
        (nz, ny, nx) or (ncomp, nz, ny, nx), or a function of the cell
        position (x, y, z)."""
        if mask is not None:
            if callable(mask):
                mask = self._eval_mask(mask)
            mask = self._check_mask_shape(_np.asarray(mask, dtype=float))
        self._impl.add_time_term(term, mask)

    def _eval_mask(self, mask):
        try:
            ncomp = len(mask(0, 0, 0))
        except TypeError:  # a scalar mask
            ncomp = 1
        X, Y, Z = self.meshgrid
        return _np.array(_np.vectorize(mask, otypes=[float] * ncomp)(X, Y, Z))

    def _check_mask_shape(self, mask):
        """(nz, ny, nx), (1, nz, ny, nx) or (ncomp, nz, ny, nx) -> (ncomp, nz, ny, nx);
        a mask with another number of components is rejected (ValueError), as
        the reference's component-wise copy rejects it (parameter.py:125-144)."""
        if mask.ndim == 4 and mask.shape[0] == self.ncomp:
            return mask
        if mask.ndim == 4 and mask.shape[0] != 1:
            raise ValueError(f"The mask has {mask.shape[0]} components; the parameter has {self.ncomp}.")
        core = mask if mask.ndim != 4 else mask[0]
        return _np.broadcast_to(core, (self.ncomp, *core.shape)).copy()

    def remove_time_terms(self):
        self._impl.remove_time_terms()

    def _reset_fields_default(self):
        self._impl.set(0.0 if self.ncomp == 1 else (0.0, 0.0, 0.0))
        self.remove_time_terms()

    def set(self, value):
        """Set a uniform value (scalar / 3-tuple), a per-cell array, a
        function of the cell position (x, y, z), a time-dependent term
        (t) -> value, or a (term, mask) pair (parameter.py:146-214 of the
        reference).  Earlier values and time terms are dropped first."""
        self._reset_fields_default()
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
        """A uniform, static value in the cells of one region (parameter.py:216-224
        of the reference)."""
        assert (isinstance(value, (float, int)) or
                (isinstance(value, tuple) and len(value) == 3)), "The value should be uniform and static."
        self._impl.set_in_region(region_idx, value)
