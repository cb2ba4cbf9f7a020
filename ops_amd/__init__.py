"""MI355X (gfx950) execution backend for the OPS structured-mesh API.

The product is the C-ABI library ops_amd/libopsmi.so (include/ops_mi355x.h);
ops_amd.ops is its Python front end.
"""
