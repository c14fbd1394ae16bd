"""CPU oracle (test infrastructure only).

Only `tests/`, `__graft_entry__.smoke()` and the CPU-baseline leg of
`bench.py` may import this package, and only as the checker.  The product
package `dalia_amd` never imports it.
"""
