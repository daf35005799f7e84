"""Board model the router and the .ORP / .ORS formats work on.

Same entities, attribute names and integrity rules as the reference's domain
model (orthoroute/domain/models/board.py: Coordinate, Bounds, Pad, Component,
Net, Layer, Board), so that boards built by the reference's loaders and by
``orthoroute_amd.serialization`` are interchangeable for
``PathFinderRouter.initialize_graph``.
"""
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple
from uuid import uuid4


@dataclass(frozen=True)
class Coordinate:
    x: float
    y: float

    def distance_to(self, other: "Coordinate") -> float:
        return ((self.x - other.x) ** 2 + (self.y - other.y) ** 2) ** 0.5


@dataclass(frozen=True)
class Bounds:
    min_x: float
    min_y: float
    max_x: float
    max_y: float

    @property
    def width(self) -> float:
        return self.max_x - self.min_x

    @property
    def height(self) -> float:
        return self.max_y - self.min_y

    @property
    def center(self) -> Coordinate:
        return Coordinate((self.min_x + self.max_x) / 2, (self.min_y + self.max_y) / 2)


@dataclass
class Pad:
    id: str
    component_id: str
    net_id: Optional[str]
    position: Coordinate
    size: Tuple[float, float]
    drill_size: Optional[float] = None
    layer: str = "F.Cu"
    shape: str = "circle"
    angle: float = 0.0

    def __post_init__(self):
        if not self.id:
            self.id = str(uuid4())


@dataclass
class Component:
    id: str
    reference: str
    value: str
    footprint: str
    position: Coordinate
    angle: float = 0.0
    layer: str = "F.Cu"
    pads: List[Pad] = field(default_factory=list)

    def __post_init__(self):
        if not self.id:
            self.id = str(uuid4())
        for pad in self.pads:
            pad.component_id = self.id

    def get_bounds(self) -> Bounds:
        """Pad rectangles' bounding box (the component position when it has
        no pads)."""
        if not self.pads:
            return Bounds(self.position.x, self.position.y, self.position.x, self.position.y)
        return Bounds(min(p.position.x - p.size[0] / 2 for p in self.pads),
                      min(p.position.y - p.size[1] / 2 for p in self.pads),
                      max(p.position.x + p.size[0] / 2 for p in self.pads),
                      max(p.position.y + p.size[1] / 2 for p in self.pads))


@dataclass
class Net:
    id: str
    name: str
    netclass: str = "Default"
    pads: List[Pad] = field(default_factory=list)

    def __post_init__(self):
        if not self.id:
            self.id = str(uuid4())
        for pad in self.pads:
            pad.net_id = self.id

    @property
    def is_routable(self) -> bool:
        return len(self.pads) >= 2

    def get_bounds(self) -> Bounds:
        if not self.pads:
            return Bounds(0, 0, 0, 0)
        xs = [p.position.x for p in self.pads]
        ys = [p.position.y for p in self.pads]
        return Bounds(min(xs), min(ys), max(xs), max(ys))


@dataclass
class Layer:
    name: str
    type: str
    stackup_position: int
    thickness: float = 0.035
    material: str = "copper"

    @property
    def is_routing_layer(self) -> bool:
        return self.type in ("signal", "power", "ground") and "Cu" in self.name


@dataclass
class Board:
    id: str
    name: str
    components: List[Component] = field(default_factory=list)
    nets: List[Net] = field(default_factory=list)
    layers: List[Layer] = field(default_factory=list)
    thickness: float = 1.6
    layer_count: int = 2
    _components_by_id: Dict[str, Component] = field(default_factory=dict, init=False, repr=False)
    _nets_by_id: Dict[str, Net] = field(default_factory=dict, init=False, repr=False)
    _nets_by_name: Dict[str, Net] = field(default_factory=dict, init=False, repr=False)

    def __post_init__(self):
        if not self.id:
            self.id = str(uuid4())
        self._components_by_id = {c.id: c for c in self.components}
        self._nets_by_id = {n.id: n for n in self.nets}
        self._nets_by_name = {n.name: n for n in self.nets}

    def get_net(self, net_id: str) -> Optional[Net]:
        return self._nets_by_id.get(net_id)

    def get_net_by_name(self, name: str) -> Optional[Net]:
        return self._nets_by_name.get(name)

    def get_component(self, component_id: str) -> Optional[Component]:
        return self._components_by_id.get(component_id)

    def get_routable_nets(self) -> List[Net]:
        return [n for n in self.nets if n.is_routable]

    def get_all_pads(self) -> List[Pad]:
        return [p for c in self.components for p in c.pads]

    def get_bounds(self) -> Bounds:
        if not self.components:
            return Bounds(0, 0, 0, 0)
        bs = [c.get_bounds() for c in self.components]
        return Bounds(min(b.min_x for b in bs), min(b.min_y for b in bs),
                      max(b.max_x for b in bs), max(b.max_y for b in bs))

    def validate_integrity(self) -> List[str]:
        """Pads naming a missing net, nets naming a missing pad, duplicate
        component references."""
        issues = []
        pads = self.get_all_pads()
        ids = {p.id for p in pads}
        for p in pads:
            if p.net_id and p.net_id not in self._nets_by_id:
                issues.append(f"Pad {p.id} references non-existent net {p.net_id}")
        for n in self.nets:
            for p in n.pads:
                if p.id not in ids:
                    issues.append(f"Net {n.name} references non-existent pad {p.id}")
        refs = [c.reference for c in self.components]
        for dup in sorted({r for r in refs if refs.count(r) > 1}):
            issues.append(f"Duplicate component reference: {dup}")
        return issues
