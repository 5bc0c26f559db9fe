"""InterParameter (mumaxplus/interparameter.py of the reference): a value
between pairs of different regions, e.g. ``inter_exchange``."""

import numpy as _np


class InterParameter:
    """A material parameter that acts between two different regions."""

    def __init__(self, impl):
        self._impl = impl

    def __repr__(self):
        return (f"InterParameter(number_of_regions={self.number_of_regions}, "
                f"name='{self.name}', ncomp={self.ncomp}, unit={self.unit})")

    def eval(self):
        """Symmetric matrix of the values: entry (i, j) holds the value
        between regions i and j (rows and columns up to the largest region
        index; the diagonal is zero)."""
        values = self._impl.eval()
        if not values:
            raise ValueError(f"The InterParameter '{self.name}' is not defined.")
        n = int(_np.max(self.region_indices)) + 1
        matrix = _np.zeros((n, n))
        i_s, j_s = _np.tril_indices(n, k=-1)
        matrix[i_s, j_s] = values
        matrix[j_s, i_s] = values
        return matrix

    def __call__(self):
        return self.eval()

    @property
    def name(self):
        return self._impl.name

    @property
    def unit(self):
        return self._impl.unit

    @property
    def ncomp(self):
        return self._impl.ncomp

    @property
    def number_of_regions(self):
        return self._impl.number_of_regions

    @property
    def region_indices(self):
        return self._impl.unique_regions

    @property
    def is_uniform(self):
        return self._impl.is_uniform

    @property
    def uniform_value(self):
        return self._impl.uniform_value

    @uniform_value.setter
    def uniform_value(self, value):
        assert isinstance(value, (float, int)), "The value should be uniform and static."
        self._impl.uniform_value = value

    def set(self, value):
        """The same value between every pair of different regions."""
        assert isinstance(value, (float, int)), "The value should be uniform and static."
        self._impl.set(value)

    def set_between(self, i, j, value):
        """The value between regions i and j."""
        self._impl.set_between(i, j, value)

    def get_between(self, i, j):
        """The value between regions i and j."""
        return self._impl.get_between(i, j)
