// Scalar types of the streaming path (same choices as the reference's
// include/types.h:39-71: dfloat=double, dlong=int, hlong=long long).
#pragma once

#define dfloat double
#define dfloatFormat "%lf"
#define dfloatString "double"

#define hlong long long int
#define hlongFormat "%lld"
#define hlongString "long long int"

#define dlong int
#define dlongFormat "%d"
#define dlongString "int"
