"""Geometry shapes: boolean functions of the cell position (x, y, z).

Semantics follow mumaxplus/util/shape.py of the reference: half-open
rectangles (shape.py:370-376), closed ellipses (x/dx)^2 + (y/dy)^2 <= 1/4
(:330-336), translation and periodic repetition (:160-192).
"""

import numpy as _np


class Shape:
    def __init__(self, shape_func=lambda x, y, z: True):
        self.shape_func = shape_func

    def __call__(self, x, y, z):
        return self.shape_func(x, y, z)

    def translate(self, dx, dy, dz):
        old = self.shape_func
        self.shape_func = lambda x, y, z: old(x - dx, y - dy, z - dz)
        return self

    def repeat(self, min_point, max_point):
        def none_mod(x, x_min, x_max):
            if x_min is None or x_max is None:
                return x
            return (x - x_min) % (x_max - x_min) + x_min

        x0, y0, z0 = min_point
        x1, y1, z1 = max_point
        old = self.shape_func
        self.shape_func = lambda x, y, z: old(none_mod(x, x0, x1), none_mod(y, y0, y1), none_mod(z, z0, z1))
        return self

    def __neg__(self):
        return Shape(lambda x, y, z: _np.logical_not(self(x, y, z)))

    def __and__(self, other):
        return Shape(lambda x, y, z: _np.logical_and(self(x, y, z), other(x, y, z)))

    def __or__(self, other):
        return Shape(lambda x, y, z: _np.logical_or(self(x, y, z), other(x, y, z)))


class Rectangle(Shape):
    """2D rectangle in the xy-plane: -sx/2 <= x < sx/2 and -sy/2 <= y < sy/2."""

    def __init__(self, sidex, sidey):
        def f(x, y, z):
            rx, ry = 0.5 * sidex, 0.5 * sidey
            return (-rx <= x) & (x < rx) & (-ry <= y) & (y < ry)
        super().__init__(f)


class Square(Rectangle):
    def __init__(self, side):
        super().__init__(side, side)


class Ellipse(Shape):
    """Ellipse in the xy-plane with diameters diamx, diamy."""

    def __init__(self, diamx, diamy):
        super().__init__(lambda x, y, z: (x / diamx) ** 2 + (y / diamy) ** 2 <= 0.25)


class Circle(Ellipse):
    def __init__(self, diam):
        super().__init__(diam, diam)
