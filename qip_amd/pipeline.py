"""Graph execution (counterpart of qip/pipeline.py).

A circuit is a graph of PipelineObject nodes built by the node constructors of ``qip_amd.qip``
and ``qip_amd.operators``. ``run`` collects the graph, gives every input node a contiguous range
of qubit indices in order of its `qid`, builds the state through the backend constructor, and
feeds every node to the state in dependency order, lowest qid first. The default backend is the
MI355X ``GpuBackend``; any StateType constructor can be passed as ``backend_constructor``.
"""
import collections
import heapq
import itertools
from typing import AbstractSet, Any, Callable, Mapping, Sequence, Tuple, Type, Union

import numpy

from .backend import GpuBackend, StateType
from .util import InitialState, flatten


class PipelineObject(object):
    """Base graph node (qip/pipeline.py:7-68)."""

    def __init__(self, quantum: bool, default: InitialState = None, n: int = 0):
        self.quantum = quantum
        self.inputs = []
        self.sink = []
        self.default = default
        self.n = n

    def run(self, feed: Mapping["PipelineObject", InitialState] = None, **kwargs):
        return run(self, feed=feed, **kwargs)

    def feed(self, state: StateType, qbitindex: Mapping["PipelineObject", Sequence[int]],
             n: int) -> Tuple[int, object]:
        """Apply this node to the state, given the qubit indices of every node."""
        return self.feed_indices(state, [qbitindex[q] for q in self.inputs], n)

    def feed_indices(self, state: StateType, index_groups: Sequence[Sequence[int]],
                     n: int) -> Tuple[int, Any]:
        """Apply this node given its inputs' index groups; returns (qubits removed, annotation)."""
        return 0, 0

    def select_index(self, indices: Sequence[int]) -> Sequence[int]:
        return indices

    def remap_index(self, index_map: Mapping["PipelineObject", int], n: int) -> Mapping["PipelineObject", int]:
        return index_map

    def set_sink(self, sink: "PipelineObject"):
        graph_changed()
        self.sink += [sink]

    def get_inputs(self) -> Sequence["PipelineObject"]:
        return self.inputs

    # Nodes compare by identity, so an identity hash is consistent with equality. (The reference
    # hashes repr(self), qip/pipeline.py:67-68, which recurses through every input and makes each
    # set operation of a run O(graph depth); its string hash is randomised per process, so no
    # result can depend on the order it gives.)
    __hash__ = object.__hash__


# Schedules of recent runs. A run's visiting order and the index groups every node is fed depend only
# on the graph (nodes, their inputs and sinks) and on the feed's keys, as long as no node overrides
# `feed` or `remap_index`; a graph run again replays its schedule instead of walking the graph. Every
# new sink (any node built on existing nodes, in any graph) advances the version and retires every
# stored schedule. A replay feeds the same nodes, in the same order, with the same index groups as
# the walk.
_GRAPH_VERSION = [0]
_SCHEDULES = collections.OrderedDict()
_SCHEDULE_SLOTS = 8


def graph_changed() -> None:
    """Retire every stored schedule (called whenever a node gains a sink)."""
    _GRAPH_VERSION[0] += 1


def _replayable(graphnodes: AbstractSet[PipelineObject]) -> bool:
    return all(type(node).feed is PipelineObject.feed and type(node).remap_index is PipelineObject.remap_index
               for node in graphnodes)


class _ScheduleRecorder:
    def __init__(self):
        self.order = []

    def feed(self, qbitindex: Mapping[PipelineObject, Sequence[int]], node: PipelineObject) -> None:
        self.order.append((node, [qbitindex[q] for q in node.inputs]))


def run(*args: PipelineObject,
        feed: Mapping[Union[PipelineObject, Tuple[PipelineObject, ...]], InitialState] = None,
        strict: bool = False,
        backend_constructor: Callable = None,
        statetype: Type = numpy.complex128,
        **kwargs):
    """Run the circuit ending in `args` (qip/pipeline.py:71-135).

    feed: {node or tuple of nodes: amplitudes}; unfed input nodes take their `default` (an int
    becomes a basis vector) or |0>. Returns (state.get_state(), {node: annotation}).
    """
    feed = {} if feed is None else {(k if type(k) == tuple else (k,)): feed[k] for k in feed}
    fed_nodes = flatten(feed.keys())
    if backend_constructor is None:
        backend_constructor = GpuBackend.make_state

    key = (args, tuple(feed.keys()))
    stored = _SCHEDULES.get(key)
    if stored is not None and stored[0] == _GRAPH_VERSION[0]:
        _SCHEDULES.move_to_end(key)
        _, frontier, graphnodes, order = stored
    else:
        frontier, graphnodes = get_deps(*args, feed=feed)
        order = None
        if _replayable(graphnodes):
            recorder = _ScheduleRecorder()
            run_graph(frontier, graphnodes, recorder)
            order = recorder.order
            _SCHEDULES[key] = (_GRAPH_VERSION[0], frontier, graphnodes, order)
            while len(_SCHEDULES) > _SCHEDULE_SLOTS:
                _SCHEDULES.popitem(last=False)
    for qubit in frontier:
        if qubit not in fed_nodes and qubit.default is not None:
            if type(qubit.default) == int:
                # default 0 is |0>, which an unfed qubit is anyway (qip/pipeline.py:103-107)
                if 0 < qubit.default < 2 ** qubit.n:
                    vec = numpy.zeros((2 ** qubit.n,))
                    vec[qubit.default] = 1.0
                    feed[(qubit,)] = vec
            else:
                feed[(qubit,)] = qubit.default

    frontier = sorted(frontier, key=lambda q: q.qid)
    if strict:
        missing = [q for q in frontier if q not in fed_nodes]
        if len(missing):
            raise ValueError("Missing Qubit states for: {}".format(missing))

    qbitindex = {}
    n = 0
    for qbit in frontier:
        qbitindex[qbit] = list(range(n, n + qbit.n))
        n += qbit.n

    feed_index_groups = [sum((qbitindex[q] for q in qs), []) for qs in feed]
    feed_states = [feed[qs] for qs in feed]
    if "state" in kwargs and kwargs["state"] is not None and _truthy(kwargs["state"]):
        # deprecated whole-state feed (qip/pipeline.py:127-131)
        feed_index_groups = [tuple(range(n))]
        feed_states = [kwargs["state"]]

    state = backend_constructor(n, feed_index_groups, feed_states, statetype=statetype)
    if order is None:
        return feed_forward(frontier, state, graphnodes)
    acc = NodeFeeder(state, n)
    for node, index_groups in order:
        acc.feed_groups(node, index_groups)
    return acc.state.get_state(), acc.classic_map


def _truthy(v) -> bool:
    # `if kwargs['state']` on an ndarray is ambiguous; the reference only uses vectors here,
    # which it tests for truth through the list/array it was given.
    if isinstance(v, numpy.ndarray):
        return v.size > 0
    return bool(v)


def get_deps(*args: PipelineObject, feed: Mapping = None) -> Tuple[Sequence[PipelineObject], AbstractSet[PipelineObject]]:
    """(input nodes, all nodes) of the graph ending in `args` (qip/pipeline.py:138-160): walk
    inputs back until a node is fed or has no inputs."""
    if feed is None:
        feed = {}
    seen = set()
    deps = set()
    todo = set(args)
    while todo:
        node = todo.pop()
        if node in feed or len(node.inputs) == 0:
            deps.add(node)
        else:
            # a node reached again through another path has been walked already (same sets as
            # walking it again, without the repeated walks of shared sub-graphs)
            todo.update(i for i in node.inputs if i not in seen)
        seen.add(node)
    return list(deps), seen


class GraphAccumulator:
    def feed(self, qbitindex: Mapping[PipelineObject, Sequence[int]], node: PipelineObject) -> None:
        raise NotImplementedError("Method not implemented.")


def run_graph(frontier: Sequence[PipelineObject], graphnodes: AbstractSet[PipelineObject], graphacc: GraphAccumulator):
    """Visit every node once its inputs are done, lowest qid first (qip/pipeline.py:186-222)."""
    qbitindex = {}
    seen = set()
    n = 0
    for qbit in sorted(frontier, key=lambda q: q.qid):
        qbitindex[qbit] = list(range(n, n + qbit.n))
        n += qbit.n
    # The reference re-sorts its frontier list by qid (a stable sort) and pops the head: lowest qid
    # first, equal qids in the order they joined the frontier. A heap keyed (qid, arrival) takes the
    # same order without re-sorting.
    arrival = itertools.count()
    heap = [(node.qid, next(arrival), node) for node in frontier]
    heapq.heapify(heap)
    queued = set(frontier)
    while heap:
        node = heapq.heappop(heap)[2]
        queued.discard(node)
        graphacc.feed(qbitindex, node)
        qbitindex = node.remap_index(qbitindex, n)
        seen.add(node)
        for nxt in node.sink:
            if nxt in graphnodes and nxt not in seen and nxt not in queued:
                if all(prev in seen for prev in nxt.inputs):
                    heapq.heappush(heap, (nxt.qid, next(arrival), nxt))
                    queued.add(nxt)
                    qbitindex[nxt] = nxt.select_index(flatten(qbitindex[j] for j in nxt.inputs))


class NodeFeeder(GraphAccumulator):
    """Feeds nodes to a state and records every node's annotation (qip/pipeline.py:225-240)."""

    def __init__(self, state: StateType, n: int):
        self.state = state
        self.n = n
        self.classic_map = {}

    def feed(self, qbitindex: Mapping[PipelineObject, Sequence[int]], node: PipelineObject):
        self._record(node, *node.feed(self.state, qbitindex, self.n))

    def feed_groups(self, node: PipelineObject, index_groups: Sequence[Sequence[int]]):
        self._record(node, *node.feed_indices(self.state, index_groups, self.n))

    def _record(self, node: PipelineObject, n_bits: int, bits: object):
        if n_bits > 0 or bits is not None:
            self.classic_map[node] = bits
            self.n -= n_bits


def feed_forward(frontier: Sequence[PipelineObject], state: StateType, graphnodes: AbstractSet[PipelineObject]):
    n = sum(f.n for f in frontier)
    acc = NodeFeeder(state, n)
    run_graph(frontier, graphnodes, acc)
    return acc.state.get_state(), acc.classic_map


def make_circuit_mat(*args: PipelineObject, backend_constructor: Callable = None) -> numpy.ndarray:
    """Dense matrix of the quantum part of a circuit, one basis state per column. The reference
    helper (qip/pipeline.py:163-178) feeds run(state=i) with an int, which its backend cannot
    index; here column i is the circuit applied to the basis vector e_i."""
    o, _ = run(*args, backend_constructor=backend_constructor)
    n = int(numpy.log2(len(o)))
    mat = numpy.zeros((2 ** n, 2 ** n), dtype=numpy.complex128)
    mat[:, 0] = o
    for i in range(1, 2 ** n):
        e = numpy.zeros(2 ** n, dtype=numpy.complex128)
        e[i] = 1.0
        o, _ = run(*args, state=e, backend_constructor=backend_constructor)
        mat[:, i] = o
    return mat
