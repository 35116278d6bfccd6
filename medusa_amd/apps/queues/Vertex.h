#pragma once
// level: BFS level (-1 unreached); mark: set by a consumer; hits: out-edges seen by a consumer
#define MEDUSA_VERTEX_FIELDS(F) F(MVT, level) F(MVT, mark) F(MVT, hits)
