"""Execution priorities of graph nodes, restated from test/src/graph/setPriorityTest.cpp:32-152.

The reference observes the order through test-only nodes (SleepNode, RecordTimeNode) and their timestamps. Here it
is read directly with rgl_graph_get_execution_order, which returns the order rgl_graph_run enqueues the nodes in
(one stream per graph, so enqueue order is execution order). Points transforms stand in for the sleep nodes and
yields for the timestamp nodes. Priority propagation to parents and the error for a parent below its children are
checked as the reference does. TwoEntryNodes (:99-146) needs a spatial merge to join two entry nodes; that node is
outside the raytrace-and-point-cloud path (rgl_node_points_spatial_merge returns RGL_NOT_IMPLEMENTED), so no graph
here has two entry nodes and that case is not restated."""
import numpy as np
import pytest

from robotecgpulidar_amd import scenes as S

pytestmark = pytest.mark.gpu

XYZ = 1


def source(rgl):
    return rgl.node_points_from_array(rgl.Node(), np.zeros((1, 3), np.float32), [XYZ])


def branch(rgl, parent):
    """parent -> 'sleep' (points transform) -> 'record' (yield); returns (sleep, record)."""
    sleep = rgl.node_points_transform(rgl.Node(), S.f32(S.identity()))
    record = rgl.node_points_yield(rgl.Node(), [XYZ])
    if parent is not None:
        rgl.add_child(parent, sleep)
    rgl.add_child(sleep, record)
    return sleep, record


def position(rgl, graph_node, node) -> int:
    return rgl.execution_order(graph_node).index(node.h.value)


def runs_in_order(rgl, entry, *nodes) -> bool:
    rgl.graph_run(entry)
    for n in nodes:  # every branch ran
        assert len(rgl.result(n, XYZ)) == 1
    pos = [position(rgl, entry, n) for n in nodes]
    return pos == sorted(pos)


def priorities_match(rgl, expected: dict) -> bool:
    return all(rgl.get_priority(n) == p for n, p in expected.items())


def test_two_branches(rgl_lib):
    rgl = rgl_lib
    src = source(rgl)
    sleep_a, node_a = branch(rgl, src)
    sleep_b, node_b = branch(rgl, src)
    order = rgl.execution_order(src)
    assert order[0] == src.h.value and len(order) == 5
    # Normally, nodes execute in the order they were added.
    assert runs_in_order(rgl, src, node_a, node_b)
    # A priority makes B run first.
    rgl.check("rgl_graph_node_set_priority", node_b.h, 1)
    assert runs_in_order(rgl, src, node_b, node_a)
    # A higher priority on A makes it run first again.
    rgl.check("rgl_graph_node_set_priority", node_a.h, 2)
    assert runs_in_order(rgl, src, node_a, node_b)
    # A parent's priority cannot drop below its children's.
    status = rgl.call("rgl_graph_node_set_priority", src.h, 1)
    assert status == rgl.RGL_INVALID_PIPELINE and "cannot set priority" in rgl.last_error()
    assert priorities_match(rgl, {node_a: 2, sleep_a: 2, src: 2, node_b: 1, sleep_b: 1})


def test_add_child(rgl_lib):
    rgl = rgl_lib
    src = source(rgl)
    sleep_a, node_a = branch(rgl, src)
    rgl.check("rgl_graph_node_set_priority", node_a.h, 1)
    # Branch B, priority -1: a lower-priority child does not lower its parent.
    sleep_b = rgl.node_points_transform(rgl.Node(), S.f32(S.identity()))
    node_b = rgl.node_points_yield(rgl.Node(), [XYZ])
    rgl.check("rgl_graph_node_set_priority", node_b.h, -1)
    rgl.add_child(sleep_b, node_b)
    assert priorities_match(rgl, {node_b: -1, sleep_b: 0})
    rgl.check("rgl_graph_node_set_priority", sleep_b.h, -1)
    # Branch C, priority 2, not yet connected.
    sleep_c, node_c = branch(rgl, None)
    rgl.check("rgl_graph_node_set_priority", node_c.h, 2)
    rgl.graph_run(src)  # once without branches B and C
    rgl.add_child(src, sleep_b)
    rgl.add_child(src, sleep_c)
    # C, A, B
    for _ in range(3):
        assert runs_in_order(rgl, src, node_c, node_a, node_b)
    assert priorities_match(rgl, {node_a: 1, sleep_a: 1, src: 2, node_b: -1, sleep_b: -1, node_c: 2, sleep_c: 2})


def test_order_follows_removal_and_readdition(rgl_lib):
    """Removing a child and adding it back appends it after its equal-priority siblings (insertion order)."""
    rgl = rgl_lib
    src = source(rgl)
    sleep_a, node_a = branch(rgl, src)
    sleep_b, node_b = branch(rgl, src)
    assert runs_in_order(rgl, src, node_a, node_b)
    rgl.check("rgl_graph_node_remove_child", src.h, sleep_a.h)
    rgl.add_child(src, sleep_a)
    assert runs_in_order(rgl, src, node_b, node_a)
