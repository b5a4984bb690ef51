"""Circuit containers of the drop-in.

Only the protocol the engine reads matters (``crates/pauli/src/engine.rs:44-70``):
``circuit.layers`` is a list of layers, each a list of rotations with
``.generator`` (PauliString or MajoranaMonomial), ``.angle`` (theta of
exp(-i theta G / 2)) and optionally ``.qiskit_gate_idx``.  Any object with
that shape works, including the reference's own circuit classes.
"""

from __future__ import annotations

from typing import List, Optional, Union


class AbstractRotation:
    def __init__(self, generator, angle: float, is_intermediate: bool = False,
                 qiskit_gate_idx: Optional[int] = None):
        self.generator = generator
        self.angle = float(angle)
        self.is_intermediate = is_intermediate
        self.qiskit_gate_idx = qiskit_gate_idx

    def __repr__(self):
        return f"{type(self).__name__}({self.generator!r}, {self.angle!r})"


class PauliRotation(AbstractRotation):
    pass


class MajoranaRotation(AbstractRotation):
    pass


class AbstractCircuit:
    def __init__(self, rotations_or_layers: Union[list, List[list]], n_modes: Optional[int] = None):
        if rotations_or_layers and isinstance(rotations_or_layers[0], list):
            self._layers = rotations_or_layers
        else:
            self._layers = [[r] for r in rotations_or_layers]
        self.n_modes = n_modes

    @property
    def layers(self):
        return self._layers

    @property
    def rotations(self):
        return [r for layer in self._layers for r in layer]

    def append(self, rotation, *, new_layer: bool = True) -> None:
        if new_layer or not self._layers:
            self._layers.append([rotation])
        else:
            self._layers[-1].append(rotation)

    def inverse(self):
        return type(self)([[type(r)(r.generator, -r.angle) for r in reversed(layer)]
                           for layer in reversed(self._layers)], self.n_modes)

    @classmethod
    def from_generators_and_angles(cls, generators, angles):
        rot = PauliRotation if cls is PauliCircuit else MajoranaRotation
        return cls([rot(g, a) for g, a in zip(generators, angles)])


class PauliCircuit(AbstractCircuit):
    pass


class MajoranaCircuit(AbstractCircuit):
    pass
