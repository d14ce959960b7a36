"""Graph execution (counterpart of qip/pipeline.py).

A circuit is a graph of PipelineObject nodes wired through two lists per node: `inputs` (the nodes
it reads) and `sink` (the nodes that registered themselves as its readers). `run` turns the part of
the graph its outputs need into a Schedule and plays it against a state backend:

  sources  the nodes the walk starts from, in qid order; each takes the next `n` qubit indices of
           the state (qip/pipeline.py:111-122). Inside run the sources are the leaves (nodes
           without inputs): run hands the collector its feed keyed by tuples, which no node equals
           (qip/pipeline.py:90,100,155).
  steps    Kahn's algorithm over the graph. A node becomes runnable when the last of its inputs to
           run lists it as a sink; of the runnable nodes the lowest qid runs first, equal qids in
           the order they became runnable (qip/pipeline.py:186-222). Each step holds the index
           groups of the node's inputs at that point of the walk.

A schedule depends only on the graph and on the feed's keys, so run keeps compiled schedules on the
first output node (at most four feed-key sets per node) and plays one again while no node anywhere
has gained a sink since it was compiled. Graphs with a node that overrides `feed` or
`remap_index` see the live index map and are walked on every run.

On a GpuBackend state of up to 256 MiB, a replayed schedule's runs of at least GRAPH_MIN_GATES plain
gate steps are captured into a HIP graph on first use and launched as one graph afterwards (a
circuit of small gates on a small state is bound by kernel launches). A graph serves a later run
while the state holds the same buffer (the engine hands a destroyed state's buffer to the next state
of its size) and every gate node of the run still holds the descriptor, and that descriptor the
kept plan, it had at capture; otherwise the run is captured again. A run that cannot be captured
(an operator whose operands do not fit kernel arguments) is applied gate by gate from then on.
"""
import heapq
import itertools
import threading
from typing import AbstractSet, Any, Callable, Dict, List, Mapping, Optional, Sequence, Tuple, Type, Union

import numpy

from .backend import GpuBackend, StateType
from .util import InitialState, flatten


class PipelineObject(object):
    """Base graph node (qip/pipeline.py:7-68). Subclasses act on the state in `feed_indices`."""

    def __init__(self, quantum: bool, default: InitialState = None, n: int = 0):
        self.quantum = quantum
        self.default = default
        self.n = n
        self.inputs = []
        self.sink = []

    def run(self, feed: Mapping["PipelineObject", InitialState] = None, **kwargs):
        return run(self, feed=feed, **kwargs)

    def feed(self, state: StateType, qbitindex: Mapping["PipelineObject", Sequence[int]],
             n: int) -> Tuple[int, object]:
        """Act on the state given every node's qubit indices; returns (qubits removed, annotation)."""
        return self.feed_indices(state, [qbitindex[q] for q in self.inputs], n)

    def feed_indices(self, state: StateType, index_groups: Sequence[Sequence[int]],
                     n: int) -> Tuple[int, Any]:
        """Act on the state given the index group of each input; the base node is the identity."""
        return 0, 0

    def select_index(self, indices: Sequence[int]) -> Sequence[int]:
        """This node's qubit indices, given its inputs' indices concatenated."""
        return indices

    def remap_index(self, index_map: Mapping["PipelineObject", int], n: int) -> Mapping["PipelineObject", int]:
        """The index map after this node has run."""
        return index_map

    def set_sink(self, sink: "PipelineObject"):
        _GRAPH_EPOCH.advance()
        self.sink.append(sink)

    def get_inputs(self) -> Sequence["PipelineObject"]:
        return self.inputs

    # Nodes compare by identity, so an identity hash is consistent with equality. (The reference
    # hashes repr(self), qip/pipeline.py:67-68, which recurses through every input and costs
    # O(graph depth) per set operation; its string hash is randomised per process, so no result
    # can depend on the order it gives.)
    __hash__ = object.__hash__


class _Epoch:
    """A counter advanced whenever any node gains a sink; compiled schedules remember it."""

    def __init__(self):
        self.value = 0

    def advance(self) -> None:
        self.value += 1


_GRAPH_EPOCH = _Epoch()
_CACHE_LOCK = threading.Lock()
GRAPH_MIN_GATES = 8          # plain gate steps in a run worth a graph
GRAPH_KEEP = 4               # graphs kept per run of steps (one per state buffer in use)
_SCHEDULES_PER_NODE = 4


def graph_changed() -> None:
    """Retire every compiled schedule (done by every set_sink; callable directly)."""
    _GRAPH_EPOCH.advance()


def _qid(node: PipelineObject) -> int:
    return getattr(node, "qid", 0)


def _collect(outputs: Sequence[PipelineObject], stop: Mapping) -> Tuple[List[PipelineObject], List[PipelineObject]]:
    """(sources, nodes) reachable backwards from `outputs`, both in depth-first discovery order:
    the walk does not go past a node without inputs or a node that is a key of `stop`."""
    nodes: List[PipelineObject] = []
    sources: List[PipelineObject] = []
    visited = set()
    pending = list(reversed(outputs))
    while pending:
        node = pending.pop()
        if node in visited:
            continue
        visited.add(node)
        nodes.append(node)
        if len(node.inputs) == 0 or node in stop:
            sources.append(node)
        else:
            pending.extend(reversed([i for i in node.inputs if i not in visited]))
    return sources, nodes


def _ordered_sources(sources: Sequence[PipelineObject]) -> List[PipelineObject]:
    # stable: equal qids keep discovery order
    return sorted(sources, key=_qid)


def _walk(sources: Sequence[PipelineObject], nodes: AbstractSet[PipelineObject],
          visit: Callable[[Dict, PipelineObject], None]) -> None:
    """Visit every node of `nodes` reachable from `sources` in execution order (module docstring).
    visit(index_map, node) sees the index map as it stands when the node runs."""
    start = _ordered_sources(sources)
    width = sum(s.n for s in start)
    index_map: Dict = {}
    base = 0
    for src in start:
        index_map[src] = list(range(base, base + src.n))
        base += src.n
    source_set = set(start)
    # readers of each node inside the graph (the inputs side of the edges); sources never wait
    readers: Dict[PipelineObject, List[PipelineObject]] = {}
    missing: Dict[PipelineObject, int] = {}
    for node in nodes:
        if node in source_set:
            continue
        distinct = set(node.inputs)
        missing[node] = len(distinct)
        for u in distinct:
            readers.setdefault(u, []).append(node)
    arrival = itertools.count()
    heap = [(_qid(s), next(arrival), s) for s in start]
    heapq.heapify(heap)
    finished = set()
    while heap:
        node = heapq.heappop(heap)[2]
        visit(index_map, node)
        index_map = node.remap_index(index_map, width)
        finished.add(node)
        now_ready = set()
        for r in readers.get(node, ()):
            missing[r] -= 1
            if missing[r] == 0:
                now_ready.add(r)
        # a node whose last input does not list it as a sink never runs (qip/pipeline.py:213-221)
        for r in node.sink:
            if r in now_ready:
                now_ready.discard(r)
                index_map[r] = r.select_index(flatten(index_map[i] for i in r.inputs))
                heapq.heappush(heap, (_qid(r), next(arrival), r))


class Schedule:
    """A graph compiled for one set of outputs and feed keys: the sources (qid order) and, when
    every node uses the base `feed` and `remap_index`, the steps as (node, index groups), the
    groups None for nodes that keep the base `feed_indices` (the identity, annotation 0)."""

    __slots__ = ("sources", "nodes", "n", "steps", "epoch", "segments", "graphs", "lock")

    def __init__(self, outputs: Sequence[PipelineObject]):
        sources, nodes = _collect(outputs, {})
        self.sources = _ordered_sources(sources)
        self.nodes = set(nodes)
        self.n = sum(s.n for s in self.sources)
        self.epoch = _GRAPH_EPOCH.value
        self.segments: Optional[Dict[int, int]] = None   # start step -> end of a run of plain gates
        self.graphs: Dict[int, Any] = {}                 # start step -> captured graphs, or False
        self.lock = threading.Lock()
        self.steps: Optional[List[Tuple[PipelineObject, Tuple[Tuple[int, ...], ...]]]] = None
        if all(type(v).feed is PipelineObject.feed and type(v).remap_index is PipelineObject.remap_index
               for v in nodes):
            steps = []

            def keep(imap, v):
                identity = type(v).feed_indices is PipelineObject.feed_indices
                steps.append((v, None if identity else tuple(tuple(imap[i]) for i in v.inputs)))
            _walk(self.sources, self.nodes, keep)
            self.steps = steps

    def index_ranges(self) -> Dict[PipelineObject, List[int]]:
        out, base = {}, 0
        for s in self.sources:
            out[s] = list(range(base, base + s.n))
            base += s.n
        return out

    def play(self, state: StateType) -> Tuple[object, Dict]:
        """Feed every node to the state; returns (state.get_state(), {node: annotation})."""
        acc = NodeFeeder(state, self.n)
        if self.steps is None:
            _walk(self.sources, self.nodes, acc.feed)
            return acc.state.get_state(), acc.classic_map
        classic = acc.classic_map
        steps = self.steps
        segments = self._segments() if getattr(state, "graphs_reusable", False) else {}
        i = 0
        while i < len(steps):
            end = segments.get(i)
            if end is not None and self._play_graph(state, i, end, acc.n):
                for node, _ in steps[i:end]:
                    classic[node] = 0        # what a plain gate step annotates
                i = end
                continue
            node, groups = steps[i]
            if groups is None:
                classic[node] = 0            # the base node: identity, annotation 0, no qubit removed
            else:
                removed, note = node.feed_indices(state, [list(g) for g in groups], acc.n)
                if removed > 0 or note is not None:
                    classic[node] = note
                    acc.n -= removed
            i += 1
        return state.get_state(), classic

    def _segments(self) -> Dict[int, int]:
        """Runs of steps that are plain gates (MatrixOp nodes with its own feed_indices) or identity
        steps, holding at least GRAPH_MIN_GATES gates: start -> end."""
        if self.segments is None:
            from .operators import MatrixOp
            plain = MatrixOp.feed_indices
            segs, start, gates = {}, None, 0
            for k, (node, groups) in enumerate(self.steps + [(None, ())]):
                ok = node is not None and (groups is None or type(node).feed_indices is plain)
                if ok:
                    if start is None:
                        start, gates = k, 0
                    gates += groups is not None
                    continue
                if start is not None and gates >= GRAPH_MIN_GATES:
                    segs[start] = k
                start = None
            self.segments = segs
        return self.segments

    def _play_graph(self, state, start: int, end: int, n: int) -> bool:
        """Apply steps[start:end] as a gate graph: a kept one that still holds, else a new capture.
        False when this run of steps cannot be captured (nothing has been applied then)."""
        gates = [(node, groups) for node, groups in self.steps[start:end] if groups is not None]
        with self.lock:
            kept = self.graphs.get(start)
        if kept is False:
            return False
        for graph, checks in kept or ():
            if state.graph_matches(graph) and all(node.ms is d and d._plan is p for node, d, p in checks):
                state.launch_graph(graph)
                return True
        state.capture_begin()
        try:
            for node, groups in gates:
                node.feed_indices(state, [list(g) for g in groups], n)
        except Exception:
            state.capture_end()          # dropped: nothing recorded has run
            with self.lock:
                self.graphs[start] = False
            return False
        graph = state.capture_end()
        if graph is None:
            with self.lock:
                self.graphs[start] = False
            return False
        checks = [(node, node.ms, getattr(node.ms, "_plan", None)) for node, _ in gates]
        state.launch_graph(graph)
        with self.lock:
            if any(p is None for _, _, p in checks):
                self.graphs[start] = False   # a descriptor without a kept plan: a change could not be seen
            else:
                kept = self.graphs.get(start) or []
                self.graphs[start] = (kept + [(graph, checks)])[-GRAPH_KEEP:]
        return True


def _schedule_for(outputs: Tuple[PipelineObject, ...], feed_keys: Tuple) -> Schedule:
    owner = outputs[0]
    key = (outputs, feed_keys)
    with _CACHE_LOCK:
        table = owner.__dict__.setdefault("_compiled_schedules", {})
        sched = table.get(key)
        if sched is not None and sched.epoch == _GRAPH_EPOCH.value:
            return sched
    sched = Schedule(outputs)
    if sched.steps is not None:
        with _CACHE_LOCK:
            table.pop(key, None)
            while len(table) >= _SCHEDULES_PER_NODE:
                table.pop(next(iter(table)))
            table[key] = sched
    return sched


def _default_vector(node: PipelineObject):
    """An unfed source's state from its `default`: an int k > 0 is the basis vector e_k, 0 is left
    out (|0> anyway), anything else is used as given (qip/pipeline.py:101-109)."""
    d = node.default
    if type(d) != int:
        return d
    if 0 < d < 2 ** node.n:
        vec = numpy.zeros((2 ** node.n,))
        vec[d] = 1.0
        return vec
    return None


def _given(v) -> bool:
    # the deprecated whole-state feed counts when non-empty (an ndarray has no plain truth value)
    if v is None:
        return False
    if isinstance(v, numpy.ndarray):
        return v.size > 0
    return bool(v)


def run(*args: PipelineObject,
        feed: Mapping[Union[PipelineObject, Tuple[PipelineObject, ...]], InitialState] = None,
        strict: bool = False,
        backend_constructor: Callable = None,
        statetype: Type = numpy.complex128,
        **kwargs):
    """Run the circuit ending in `args` (qip/pipeline.py:71-135).

    feed: {node or tuple of nodes: amplitudes}; a tuple key feeds the joint state of its nodes.
    Unfed sources take their `default` or |0>. strict=True requires every source in the feed.
    Returns (state.get_state(), {node: annotation}) for every node that ran.
    """
    given = {}
    for k, v in (feed or {}).items():
        given[k if type(k) == tuple else (k,)] = v
    fed = set(flatten(given.keys()))
    sched = _schedule_for(tuple(args), tuple(given.keys()))
    if strict:
        missing = [s for s in sched.sources if s not in fed]
        if missing:
            raise ValueError("Missing Qubit states for: {}".format(missing))
    for s in sched.sources:
        if s not in fed and s.default is not None:
            vec = _default_vector(s)
            if vec is not None:
                given[(s,)] = vec
    ranges = sched.index_ranges()
    groups = [[i for node in key for i in ranges[node]] for key in given]
    vectors = list(given.values())
    if _given(kwargs.get("state")):
        # deprecated: one vector for the whole register (qip/pipeline.py:127-131)
        groups = [tuple(range(sched.n))]
        vectors = [kwargs["state"]]
    if backend_constructor is None:
        backend_constructor = GpuBackend.make_state
    state = backend_constructor(sched.n, groups, vectors, statetype=statetype)
    return sched.play(state)


def get_deps(*args: PipelineObject, feed: Mapping = None) -> Tuple[Sequence[PipelineObject], AbstractSet[PipelineObject]]:
    """(sources, all nodes) of the graph ending in `args` (qip/pipeline.py:138-160): the walk back
    through inputs stops at nodes without inputs and at keys of `feed`."""
    sources, nodes = _collect(args, feed or {})
    return sources, set(nodes)


class GraphAccumulator:
    """Receives every node of a walk in execution order (qip/pipeline.py:181-183)."""

    def feed(self, qbitindex: Mapping[PipelineObject, Sequence[int]], node: PipelineObject) -> None:
        raise NotImplementedError("Method not implemented.")


def run_graph(frontier: Sequence[PipelineObject], graphnodes: AbstractSet[PipelineObject], graphacc: GraphAccumulator):
    """Pass every node to graphacc.feed(index_map, node) in execution order (qip/pipeline.py:186-222)."""
    _walk(list(frontier), set(graphnodes), lambda imap, node: graphacc.feed(imap, node))


class NodeFeeder(GraphAccumulator):
    """Feeds nodes to a state and keeps each node's annotation (qip/pipeline.py:225-240): every node
    whose annotation is not None lands in classic_map, 0 for plain gates."""

    def __init__(self, state: StateType, n: int):
        self.state = state
        self.n = n
        self.classic_map = {}

    def feed(self, qbitindex: Mapping[PipelineObject, Sequence[int]], node: PipelineObject):
        self._keep(node, node.feed(self.state, qbitindex, self.n))

    def feed_groups(self, node: PipelineObject, index_groups: Sequence[Sequence[int]]):
        self._keep(node, node.feed_indices(self.state, index_groups, self.n))

    def _keep(self, node: PipelineObject, result: Tuple[int, object]) -> None:
        removed, note = result
        if removed > 0 or note is not None:
            self.classic_map[node] = note
            self.n -= removed


def feed_forward(frontier: Sequence[PipelineObject], state: StateType, graphnodes: AbstractSet[PipelineObject]):
    acc = NodeFeeder(state, sum(f.n for f in frontier))
    run_graph(frontier, graphnodes, acc)
    return acc.state.get_state(), acc.classic_map


def make_circuit_mat(*args: PipelineObject, backend_constructor: Callable = None) -> numpy.ndarray:
    """Dense matrix of the quantum part of a circuit: column i is the circuit applied to the basis
    vector e_i. (The reference helper, qip/pipeline.py:163-178, feeds run(state=i) with an int,
    which its backend cannot index.)"""
    first, _ = run(*args, backend_constructor=backend_constructor)
    dim = len(first)
    cols = [numpy.asarray(first, dtype=numpy.complex128)]
    for i in range(1, dim):
        e = numpy.zeros(dim, dtype=numpy.complex128)
        e[i] = 1.0
        out, _ = run(*args, state=e, backend_constructor=backend_constructor)
        cols.append(numpy.asarray(out, dtype=numpy.complex128))
    return numpy.stack(cols, axis=1)
