// Multi-source BFS: one bit per source in a 64-bit mask.
#pragma once
#define MVT unsigned long long
#define MVT_Init_Value 0
