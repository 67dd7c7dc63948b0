"""ORACLE -- test infrastructure: CPU restatement of the reference's bitstream
semantics used as the checker for the HIP path.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import it."""
