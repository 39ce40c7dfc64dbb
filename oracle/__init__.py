"""Oracle: CPU restatement of the reference H8 path (TEST INFRASTRUCTURE ONLY).

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
as the checker.  The shipped package never imports anything from here.
"""
