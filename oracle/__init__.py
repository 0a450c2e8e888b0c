"""CPU oracle for the pRRTC planner path -- test infrastructure only (tests/, smoke(), bench.py's
cpu_baseline).  The product package never imports it."""
