// SSSP distances as floats (exact for integer weights while sums stay below 2^24).
#pragma once
#define MVT float
#define MVT_Init_Value -1
