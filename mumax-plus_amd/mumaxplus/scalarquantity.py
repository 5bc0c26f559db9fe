"""ScalarQuantity (mumaxplus/scalarquantity.py of the reference)."""


class ScalarQuantity:
    """A functor representing a physical scalar quantity."""

    def __init__(self, impl):
        self._impl = impl

    def __repr__(self):
        return f"ScalarQuantity(name='{self.name}', unit={self.unit})"

    @property
    def name(self):
        return self._impl.name

    @property
    def unit(self):
        return self._impl.unit

    def eval(self):
        return self._impl.eval()

    def __call__(self):
        return self.eval()
