"""PlatoEngine operation surface of the analyze performer (MPMD_App, src/Analyze_App.hpp
and src/Analyze_App.cpp): named operations over one elliptic problem, fields imported and
exported by name.

* Problem: an input deck (`deck.ElasticProblem`, the "Plato Problem" list of the app's
  input parameters) and its mesh (`exodus.read_exodus` of "Input Mesh", or a mesh object).
* Operations (src/Analyze_App.cpp:350-415): each `<Operation>` of the app file has a
  "Function", a "Name" and, for criterion operations, "Criterion", optional "Target" and
  `<Output>` nodes mapping an "Argument" (Value, Gradient) to an "ArgumentName"
  (src/Analyze_App.cpp:81-118, 563-576).  Supported functions: ComputeSolution,
  ComputeCriterion, ComputeCriterionValue, ComputeCriterionGradient, ComputeCriterionX,
  ComputeCriterionGradientX, UpdateProblem, Reinitialize, WriteOutput (a no-op there too).
* Data: `importData("Topology", z)` sets the control (src/Analyze_App.hpp:196-230);
  `exportData(name)` returns a criterion value (float) or gradient (device tensor) by its
  ArgumentName, or the state for "Solution".

The control starts at 1 on every node, as MPMD_App::initialize (src/Analyze_App.cpp:320-333).
"""
from __future__ import annotations

import xml.etree.ElementTree as ET
from dataclasses import dataclass, field
from pathlib import Path

import torch

from .deck import ElasticProblem, ParameterList, read_parameter_list
from .exodus import ExodusMesh, read_exodus

CRITERION_FUNCTIONS = ("ComputeCriterion", "ComputeCriterionValue", "ComputeCriterionGradient",
                       "ComputeCriterionX", "ComputeCriterionGradientX")
FUNCTIONS = CRITERION_FUNCTIONS + ("ComputeSolution", "UpdateProblem", "Reinitialize", "WriteOutput")


@dataclass
class Operation:
    function: str
    name: str
    criterion: str = ""
    target: float = 0.0
    outputs: dict = field(default_factory=dict)  # Argument -> ArgumentName


def _text(node, tag: str, default: str | None = None) -> str:
    el = node.find(tag)
    if el is None or el.text is None:
        if default is None:
            raise ValueError(f"<{node.tag}>: required keyword <{tag}> missing")
        return default
    return el.text.strip()


def parse_operations(text: str) -> list[Operation]:
    """The `<Operation>` nodes of an analyze app file."""
    root = ET.fromstring(text)
    ops = []
    for node in root.iter("Operation"):
        op = Operation(_text(node, "Function"), _text(node, "Name"))
        if op.function not in FUNCTIONS:
            raise ValueError(f"operation {op.name!r}: function {op.function!r} is not supported ({FUNCTIONS})")
        if op.function in CRITERION_FUNCTIONS:
            op.criterion = _text(node, "Criterion", "")
            if not op.criterion:
                raise ValueError("Required keyword ('Criterion') missing.")
            op.target = float(_text(node, "Target", "0.0"))
            for out in node.findall("Output"):
                arg, name = _text(out, "Argument", ""), _text(out, "ArgumentName", "")
                if arg and not name:
                    raise ValueError(f"operation {op.name!r}: 'ArgumentName' missing for {arg!r}")
                if arg:
                    op.outputs[arg] = name
        ops.append(op)
    return ops


class MPMDApp:
    """MPMD_App: compute(operation name), importData / exportData by argument name."""

    def __init__(self, params: ParameterList | str | Path, operations: list[Operation] | str | Path,
                 mesh: ExodusMesh | None = None, device: str = "cuda"):
        if not isinstance(params, ParameterList):
            params = read_parameter_list(params)
        if not isinstance(operations, list):
            p = Path(operations)
            operations = parse_operations(p.read_text() if p.exists() else str(operations))
        if mesh is None:
            if not params.isParameter("Input Mesh"):
                raise ValueError("Analyze Application: 'Input Mesh' keyword was not defined. Use the 'Input Mesh' "
                                 "keyword to provide the name of the mesh file.")
            mesh = read_exodus(params.get("Input Mesh"))
        dim = int(params.get("Spatial Dimension", 3))
        if dim != 3:
            raise ValueError(f"only 3-D problems are supported, 'Spatial Dimension' is {dim}")
        self.params, self.mesh, self.device = params, mesh, device
        self.operations: dict[str, Operation] = {}
        self.value_names: dict[str, str] = {}      # ArgumentName -> criterion (values)
        self.grad_z_names: dict[str, str] = {}     # ArgumentName -> criterion (dF/dz)
        self.grad_x_names: dict[str, str] = {}     # ArgumentName -> criterion (dF/dX)
        for op in operations:
            if op.name in self.operations:
                raise ValueError(f"operation name {op.name!r} is defined twice")
            self.operations[op.name] = op
            self._register(op)
        self._create()

    def _register(self, op: Operation) -> None:
        def add(table, arg):
            if arg in op.outputs:
                name = op.outputs[arg]
                if table.get(name, op.criterion) != op.criterion:
                    raise ValueError(f"ArgumentName {name!r} is already associated with criterion {table[name]!r}")
                table[name] = op.criterion
        if op.function in ("ComputeCriterion", "ComputeCriterionValue", "ComputeCriterionX"):
            add(self.value_names, "Value")
        if op.function in ("ComputeCriterion", "ComputeCriterionGradient"):
            add(self.grad_z_names, "Gradient")
        if op.function in ("ComputeCriterionX", "ComputeCriterionGradientX"):
            add(self.grad_x_names, "Gradient")

    def _create(self) -> None:
        self.problem = ElasticProblem(self.params, self.mesh, self.device)
        for op in self.operations.values():
            if op.criterion:
                self.problem.criterion(op.criterion)  # unknown criteria fail at set-up
        n = self.mesh.mesh.num_nodes
        self.control = torch.ones(n, dtype=torch.float64, device=self.device)
        self.state = torch.zeros(3 * n, dtype=torch.float64, device=self.device)
        self.values: dict[str, float] = {}
        self.grad_z: dict[str, torch.Tensor] = {}
        self.grad_x: dict[str, torch.Tensor] = {}

    # -------------------------------------------------------------- data exchange
    def importData(self, name: str, data) -> None:  # noqa: N802 - reference name
        if name == "Topology":
            z = torch.as_tensor(data, dtype=torch.float64).to(self.device).reshape(-1)
            if z.numel() != self.control.numel():
                raise ValueError(f"Topology has {z.numel()} values, the mesh {self.control.numel()} nodes")
            self.control.copy_(z)
        elif name == "Solution":
            u = torch.as_tensor(data, dtype=torch.float64).to(self.device).reshape(-1)
            if u.numel() != self.state.numel():
                raise ValueError(f"Solution has {u.numel()} values, expected {self.state.numel()}")
            self.state.copy_(u)
        else:
            raise KeyError(f"importData: unknown field {name!r} (Topology, Solution)")

    def exportData(self, name: str):  # noqa: N802
        if name == "Solution":
            return self.state
        if name in self.value_names:
            return self.values[self.value_names[name]]
        if name in self.grad_z_names:
            return self.grad_z[self.grad_z_names[name]]
        if name in self.grad_x_names:
            return self.grad_x[self.grad_x_names[name]]
        raise KeyError(f"exportData: no operation outputs {name!r}")

    # -------------------------------------------------------------- operations
    def compute(self, operation: str) -> None:
        if operation not in self.operations:
            raise KeyError(f"Request for operation ('{operation}') that doesn't exist.")
        op = self.operations[operation]
        pb, z = self.problem, self.control
        f = op.function
        if f == "ComputeSolution":
            self.state = pb.solution(z)
        elif f == "UpdateProblem":
            pb.updateProblem(z, self.state)
        elif f == "Reinitialize":
            z0 = self.control.clone()
            self._create()
            self.control.copy_(z0)
        elif f == "WriteOutput":
            pass
        else:
            c = op.criterion
            if f != "ComputeCriterionGradient" and f != "ComputeCriterionGradientX" and not pb.criterionIsLinear(c):
                self.state = pb.solution(z)
            if f in ("ComputeCriterion", "ComputeCriterionValue", "ComputeCriterionX"):
                self.values[c] = pb.criterionValue(z, self.state, c) - op.target
            if f in ("ComputeCriterion", "ComputeCriterionGradient"):
                self.grad_z[c] = pb.criterionGradient(z, self.state, c)
            if f in ("ComputeCriterionX", "ComputeCriterionGradientX"):
                self.grad_x[c] = pb.criterionGradientX(z, self.state, c)
