#pragma once
// edges carry no user data for BFS
#define MEDUSA_EDGE_FIELDS(F)
