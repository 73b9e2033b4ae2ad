"""Sibling process modules of the transient path (same initialize/update/
finalize contract as IGM's processes): thk, time, smb_simple."""
