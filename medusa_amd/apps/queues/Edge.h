#pragma once
// `weight` is filled from the graph's arc weights
#define MEDUSA_EDGE_FIELDS(F) F(MVT, weight)
