"""DmiTensor (mumaxplus/dmitensor.py of the reference)."""

from .parameter import Parameter


class DmiTensor:
    """The nine DMI parameters D_ijk (antisymmetric in j, k)."""

    KEYS = ("xxy", "xyz", "xxz", "yxy", "yyz", "yxz", "zxy", "zyz", "zxz")

    def __init__(self, impl):
        object.__setattr__(self, "_impl", impl)

    def __getattr__(self, key):
        if key in DmiTensor.KEYS:
            return Parameter(getattr(object.__getattribute__(self, "_impl"), key))
        raise AttributeError(key)

    def __setattr__(self, key, value):
        if key in DmiTensor.KEYS:
            getattr(self, key).set(value)
        else:
            object.__setattr__(self, key, value)

    def make_zero(self):
        for k in DmiTensor.KEYS:
            getattr(self, k).set(0.0)

    def set_interfacial_dmi(self, value):
        """xxz = yyz = value (dmitensor.py:106-124 of the reference)."""
        self.make_zero()
        self.xxz.set(value)
        self.yyz.set(value)

    def set_bulk_dmi(self, value):
        """xyz = value, yxz = -value, zxy = value (dmitensor.py:126-147)."""
        self.make_zero()
        self.xyz.set(value)
        self.yxz.set(-value)
        self.zxy.set(value)


class DmiTensorGroup:
    """Several DMI tensors set together (dmitensor.py:5-29 of the reference):
    element assignment and the set_* methods apply to every tensor."""

    def __init__(self, tensors):
        object.__setattr__(self, "tensors", list(tensors))

    def __getattr__(self, name):
        if name in ("make_zero", "set_interfacial_dmi", "set_bulk_dmi"):
            def method(*args, **kwargs):
                for tensor in self.tensors:
                    getattr(tensor, name)(*args, **kwargs)
            return method
        raise AttributeError(f"'DmiTensor' object has no attribute '{name}'")

    def __setattr__(self, name, value):
        if name in DmiTensor.KEYS:
            for tensor in self.tensors:
                setattr(tensor, name, value)
        else:
            raise AttributeError(f"'DmiTensor' object has no attribute '{name}'")
