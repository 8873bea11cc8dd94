"""``launch(kernel, dim, inputs)``: the wp.launch call shape the reference uses
at example_fluid.py:90-93,136-138, dispatching to the HIP C-ABI."""
from __future__ import annotations


class Kernel:
    """A named C-ABI entry point callable as launch(kernel, dim=..., inputs=[...])."""

    def __init__(self, name, fn, dim_of):
        self.name = name
        self._fn = fn
        self._dim_of = dim_of  # inputs -> expected launch dim

    def launch(self, dim, inputs):
        expected = tuple(self._dim_of(*inputs))
        got = tuple(dim) if not isinstance(dim, int) else (dim,)
        if got != expected:
            raise ValueError(f"{self.name}: launch dim {got} does not match the inputs ({expected})")
        self._fn(*inputs)

    def __repr__(self):
        return f"<kernel {self.name}>"


def launch(kernel: Kernel, dim, inputs, device=None, stream=None):
    kernel.launch(dim, list(inputs))
