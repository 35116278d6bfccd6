#pragma once
// one user field per vertex: its BFS level
#define MEDUSA_VERTEX_FIELDS(F) F(MVT, level)
