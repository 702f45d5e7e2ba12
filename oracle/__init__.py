"""CPU oracle of the streaming hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
this package, and only to check results against (or as the timed CPU baseline).  The
product library never imports it.
"""
