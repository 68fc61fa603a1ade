"""The benchmark queries as physical plans, runnable on any engine.

Each query is written the way MiniLake's planner would lay it out
(crates/minilake-exec/src/executor.rs:101-249): builds first, then the probe
pipeline into the aggregate, then the TopN. Build sides are the smaller
estimated input (crates/minilake-sql/src/physical_planner.rs:127-129) and
pruned to the columns used downstream.

An *engine* provides ``table``, ``join_build``, ``hashagg``, ``topn`` (and the
single-operator entry points); the CPU oracle and the GPU library both do.
"""

from __future__ import annotations

from .abi import (AGG_AVG, AGG_SUM, DATE, FLOAT64, INT32, INT64, UTF8, Agg, col, count_star, date_lit,
                  filter_op, lit, probe_op, project_op)
from . import workloads as W


def q1_plan():
    """filter + low-cardinality GROUP BY over lineitem (queries/tpch/q1.sql:1-21).

    Columns: 0 qty, 1 price, 2 disc, 3 tax, 4 flag1, 5 flag2, 6 date. The
    literal 1 in ``1 - disc`` is coerced to Float64 (binder.rs:934-953).
    """
    price, disc, tax = col(1), col(2), col(3)
    disc_price = price * (lit(1.0) - disc)
    charge = disc_price * (lit(1.0) + tax)
    ops = [filter_op(col(6).le(date_lit(W.FILTER_GROUPBY_CUTOFF)))]
    group = [col(4), col(5)]
    key_types = [UTF8, UTF8]
    aggs = [
        Agg(AGG_SUM, col(0), INT32),
        Agg(AGG_SUM, price, FLOAT64),
        Agg(AGG_SUM, disc_price, FLOAT64),
        Agg(AGG_SUM, charge, FLOAT64),
        Agg(AGG_AVG, col(0), INT32),
        Agg(AGG_AVG, price, FLOAT64),
        Agg(AGG_AVG, disc, FLOAT64),
        count_star(),
    ]
    return ops, group, key_types, aggs


def run_q1(engine, lineitem, **agg_kw):
    ops, group, kt, aggs = q1_plan()
    return engine.hashagg(lineitem, ops, group, kt, aggs, **agg_kw)


def run_q3(engine, customer, orders, lineitem, limit=10, **agg_kw):
    """Two-way join + (orderkey, orderdate, shippriority) GROUP BY + top-10 (q3.sql:1-23)."""
    cust = engine.join_build(customer, [filter_op(col(1).eq(lit("BUILDING"))), project_op([col(0)])],
                             [col(0)], [INT64])
    # orders -> probe(customer) ; layout [c_custkey, o_orderkey, o_custkey, o_orderdate, o_shippriority]
    ords = engine.join_build(orders, [filter_op(col(2).lt(date_lit(W.D_1995_03_15))),
                                      probe_op(cust, [col(1)], build_is_left=True),
                                      project_op([col(1), col(3), col(4)])],
                             [col(0)], [INT64])
    # lineitem -> probe(orders) ; layout [o_orderkey, o_orderdate, o_shippriority, l_orderkey, price, disc, ship]
    agg = engine.hashagg(lineitem,
                         [filter_op(col(3).gt(date_lit(W.D_1995_03_15))),
                          probe_op(ords, [col(0)], build_is_left=True)],
                         [col(3), col(1), col(2)], [INT64, DATE, INT32],
                         [Agg(AGG_SUM, col(4) * (lit(1.0) - col(5)), FLOAT64)], **agg_kw)
    # agg layout [l_orderkey, o_orderdate, o_shippriority, revenue]
    top = engine.topn(agg, [], [(col(3), False, False), (col(1), True, False)], limit)
    return top, (cust, ords, agg)


def run_q5(engine, sj, **agg_kw):
    """Six-table star join GROUP BY n_name (q5.sql:1-24); c_nationkey = s_nationkey is a join key."""
    # the four independent build sides together, then orders (which probes customer)
    cust, supp, nation, region = engine.join_build_many([
        (sj.customer, [], [col(0)], [INT64]),
        (sj.supplier, [], [col(0), col(1)], [INT64, INT32]),
        (sj.nation, [], [col(0)], [INT32]),
        (sj.region, [filter_op(col(1).eq(lit("ASIA")))], [col(0)], [INT32])])
    # orders -> probe(customer): [c_custkey, c_nationkey, o_orderkey, o_custkey, o_orderdate]
    ords = engine.join_build(sj.orders,
                             [filter_op(col(2).ge(date_lit(W.D_1994_01_01)) & col(2).lt(date_lit(W.D_1995_01_01))),
                              probe_op(cust, [col(1)], build_is_left=True),
                              project_op([col(1), col(2)])],
                             [col(1)], [INT64])
    # lineitem [l_orderkey, l_suppkey, price, disc]
    #  -> probe(ords, build left)  : [c_nationkey, o_orderkey, l_orderkey, l_suppkey, price, disc]
    #  -> probe(supp)              : [..6.., s_suppkey, s_nationkey]
    #  -> probe(nation)            : [..8.., n_nationkey, n_name, n_regionkey]
    #  -> probe(region)            : [..11.., r_regionkey, r_name]
    agg = engine.hashagg(sj.lineitem,
                         [probe_op(ords, [col(0)], build_is_left=True),
                          probe_op(supp, [col(3), col(0)]),
                          probe_op(nation, [col(7)]),
                          probe_op(region, [col(10)])],
                         [col(9)], [UTF8], [Agg(AGG_SUM, col(4) * (lit(1.0) - col(5)), FLOAT64)], **agg_kw)
    top = engine.topn(agg, [], [(col(1), False, False)], 25)
    return top, (cust, ords, supp, nation, region, agg)
