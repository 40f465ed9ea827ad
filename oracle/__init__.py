"""CPU oracle of the stockholder partitioning loop (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this
package, as the checker or as the timed CPU baseline; the product package
never imports it.
"""
