// bc_libm.h — a host/device double exp that returns the same bits as the x86-64 glibc 2.35
// routine the reference CPU tick calls (::exp), and correctly rounded division helpers.
//
// exp follows the published table-driven algorithm of glibc's exp (ARM optimized-routines
// family, N = 128 table, degree-5 polynomial), with the FMA contractions of glibc's
// x86-64 FMA variant written out explicitly. The table below was generated from
// 2^(k/128) evaluated in 80-digit decimal arithmetic, H[k] = nearest double,
// T[k] = 2^(k/128)/H[k] - 1 rounded to double, stored as (bits(T), bits(H) - k<<45).
// The bit-identity with the host libm is checked by tests/test_libm_replica.py on
// tens of millions of inputs; the GPU runs the same functions.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define BC_HD __host__ __device__ static inline

namespace bc {

__device__ __constant__ static const uint64_t kExpTabDev[256] = {
0x0000000000000000ULL, 0x3ff0000000000000ULL,
0x3c9b3b4f1a88bf6eULL, 0x3feff63da9fb3335ULL,
0xbc7160139cd8dc5dULL, 0x3fefec9a3e778061ULL,
0xbc905e7a108766d1ULL, 0x3fefe315e86e7f85ULL,
0x3c8cd2523567f613ULL, 0x3fefd9b0d3158574ULL,
0xbc8bce8023f98efaULL, 0x3fefd06b29ddf6deULL,
0x3c60f74e61e6c861ULL, 0x3fefc74518759bc8ULL,
0x3c90a3e45b33d399ULL, 0x3fefbe3ecac6f383ULL,
0x3c979aa65d837b6dULL, 0x3fefb5586cf9890fULL,
0x3c8eb51a92fdeffcULL, 0x3fefac922b7247f7ULL,
0x3c3ebe3d702f9cd1ULL, 0x3fefa3ec32d3d1a2ULL,
0xbc6a033489906e0bULL, 0x3fef9b66affed31bULL,
0xbc9556522a2fbd0eULL, 0x3fef9301d0125b51ULL,
0xbc5080ef8c4eea55ULL, 0x3fef8abdc06c31ccULL,
0xbc91c923b9d5f416ULL, 0x3fef829aaea92de0ULL,
0x3c80d3e3e95c55afULL, 0x3fef7a98c8a58e51ULL,
0xbc801b15eaa59348ULL, 0x3fef72b83c7d517bULL,
0xbc8f1ff055de323dULL, 0x3fef6af9388c8deaULL,
0x3c8b898c3f1353bfULL, 0x3fef635beb6fcb75ULL,
0xbc96d99c7611eb26ULL, 0x3fef5be084045cd4ULL,
0x3c9aecf73e3a2f60ULL, 0x3fef54873168b9aaULL,
0xbc8fe782cb86389dULL, 0x3fef4d5022fcd91dULL,
0x3c8a6f4144a6c38dULL, 0x3fef463b88628cd6ULL,
0x3c807a05b0e4047dULL, 0x3fef3f49917ddc96ULL,
0x3c968efde3a8a894ULL, 0x3fef387a6e756238ULL,
0x3c875e18f274487dULL, 0x3fef31ce4fb2a63fULL,
0x3c80472b981fe7f2ULL, 0x3fef2b4565e27cddULL,
0xbc96b87b3f71085eULL, 0x3fef24dfe1f56381ULL,
0x3c82f7e16d09ab31ULL, 0x3fef1e9df51fdee1ULL,
0xbc3d219b1a6fbffaULL, 0x3fef187fd0dad990ULL,
0x3c8b3782720c0ab4ULL, 0x3fef1285a6e4030bULL,
0x3c6e149289cecb8fULL, 0x3fef0cafa93e2f56ULL,
0x3c834d754db0abb6ULL, 0x3fef06fe0a31b715ULL,
0x3c864201e2ac744cULL, 0x3fef0170fc4cd831ULL,
0x3c8fdd395dd3f84aULL, 0x3feefc08b26416ffULL,
0xbc86a3803b8e5b04ULL, 0x3feef6c55f929ff1ULL,
0xbc924aedcc4b5068ULL, 0x3feef1a7373aa9cbULL,
0xbc9907f81b512d8eULL, 0x3feeecae6d05d866ULL,
0xbc71d1e83e9436d2ULL, 0x3feee7db34e59ff7ULL,
0xbc991919b3ce1b15ULL, 0x3feee32dc313a8e5ULL,
0x3c859f48a72a4c6dULL, 0x3feedea64c123422ULL,
0xbc9312607a28698aULL, 0x3feeda4504ac801cULL,
0xbc58a78f4817895bULL, 0x3feed60a21f72e2aULL,
0xbc7c2c9b67499a1bULL, 0x3feed1f5d950a897ULL,
0x3c4363ed60c2ac11ULL, 0x3feece086061892dULL,
0x3c9666093b0664efULL, 0x3feeca41ed1d0057ULL,
0x3c6ecce1daa10379ULL, 0x3feec6a2b5c13cd0ULL,
0x3c93ff8e3f0f1230ULL, 0x3feec32af0d7d3deULL,
0x3c7690cebb7aafb0ULL, 0x3feebfdad5362a27ULL,
0x3c931dbdeb54e077ULL, 0x3feebcb299fddd0dULL,
0xbc8f94340071a38eULL, 0x3feeb9b2769d2ca7ULL,
0xbc87deccdc93a349ULL, 0x3feeb6daa2cf6642ULL,
0xbc78dec6bd0f385fULL, 0x3feeb42b569d4f82ULL,
0xbc861246ec7b5cf6ULL, 0x3feeb1a4ca5d920fULL,
0x3c93350518fdd78eULL, 0x3feeaf4736b527daULL,
0x3c7b98b72f8a9b05ULL, 0x3feead12d497c7fdULL,
0x3c9063e1e21c5409ULL, 0x3feeab07dd485429ULL,
0x3c34c7855019c6eaULL, 0x3feea9268a5946b7ULL,
0x3c9432e62b64c035ULL, 0x3feea76f15ad2148ULL,
0xbc8ce44a6199769fULL, 0x3feea5e1b976dc09ULL,
0xbc8c33c53bef4da8ULL, 0x3feea47eb03a5585ULL,
0xbc845378892be9aeULL, 0x3feea34634ccc320ULL,
0xbc93cedd78565858ULL, 0x3feea23882552225ULL,
0x3c5710aa807e1964ULL, 0x3feea155d44ca973ULL,
0xbc93b3efbf5e2228ULL, 0x3feea09e667f3bcdULL,
0xbc6a12ad8734b982ULL, 0x3feea012750bdabfULL,
0xbc6367efb86da9eeULL, 0x3fee9fb23c651a2fULL,
0xbc80dc3d54e08851ULL, 0x3fee9f7df9519484ULL,
0xbc781f647e5a3ecfULL, 0x3fee9f75e8ec5f74ULL,
0xbc86ee4ac08b7db0ULL, 0x3fee9f9a48a58174ULL,
0xbc8619321e55e68aULL, 0x3fee9feb564267c9ULL,
0x3c909ccb5e09d4d3ULL, 0x3feea0694fde5d3fULL,
0xbc7b32dcb94da51dULL, 0x3feea11473eb0187ULL,
0x3c94ecfd5467c06bULL, 0x3feea1ed0130c132ULL,
0x3c65ebe1abd66c55ULL, 0x3feea2f336cf4e62ULL,
0xbc88a1c52fb3cf42ULL, 0x3feea427543e1a12ULL,
0xbc9369b6f13b3734ULL, 0x3feea589994cce13ULL,
0xbc805e843a19ff1eULL, 0x3feea71a4623c7adULL,
0xbc94d450d872576eULL, 0x3feea8d99b4492edULL,
0x3c90ad675b0e8a00ULL, 0x3feeaac7d98a6699ULL,
0x3c8db72fc1f0eab4ULL, 0x3feeace5422aa0dbULL,
0xbc65b6609cc5e7ffULL, 0x3feeaf3216b5448cULL,
0x3c7bf68359f35f44ULL, 0x3feeb1ae99157736ULL,
0xbc93091fa71e3d83ULL, 0x3feeb45b0b91ffc6ULL,
0xbc5da9b88b6c1e29ULL, 0x3feeb737b0cdc5e5ULL,
0xbc6c23f97c90b959ULL, 0x3feeba44cbc8520fULL,
0xbc92434322f4f9aaULL, 0x3feebd829fde4e50ULL,
0xbc85ca6cd7668e4bULL, 0x3feec0f170ca07baULL,
0x3c71affc2b91ce27ULL, 0x3feec49182a3f090ULL,
0x3c6dd235e10a73bbULL, 0x3feec86319e32323ULL,
0xbc87c50422622263ULL, 0x3feecc667b5de565ULL,
0x3c8b1c86e3e231d5ULL, 0x3feed09bec4a2d33ULL,
0xbc91bbd1d3bcbb15ULL, 0x3feed503b23e255dULL,
0x3c90cc319cee31d2ULL, 0x3feed99e1330b358ULL,
0x3c8469846e735ab3ULL, 0x3feede6b5579fdbfULL,
0xbc82dfcd978e9db4ULL, 0x3feee36bbfd3f37aULL,
0x3c8c1a7792cb3387ULL, 0x3feee89f995ad3adULL,
0xbc907b8f4ad1d9faULL, 0x3feeee07298db666ULL,
0xbc55c3d956dcaebaULL, 0x3feef3a2b84f15fbULL,
0xbc90a40e3da6f640ULL, 0x3feef9728de5593aULL,
0xbc68d6f438ad9334ULL, 0x3feeff76f2fb5e47ULL,
0xbc91eee26b588a35ULL, 0x3fef05b030a1064aULL,
0x3c74ffd70a5fddcdULL, 0x3fef0c1e904bc1d2ULL,
0xbc91bdfbfa9298acULL, 0x3fef12c25bd71e09ULL,
0x3c736eae30af0cb3ULL, 0x3fef199bdd85529cULL,
0x3c8ee3325c9ffd94ULL, 0x3fef20ab5fffd07aULL,
0x3c84e08fd10959acULL, 0x3fef27f12e57d14bULL,
0x3c63cdaf384e1a67ULL, 0x3fef2f6d9406e7b5ULL,
0x3c676b2c6c921968ULL, 0x3fef3720dcef9069ULL,
0xbc808a1883ccb5d2ULL, 0x3fef3f0b555dc3faULL,
0xbc8fad5d3ffffa6fULL, 0x3fef472d4a07897cULL,
0xbc900dae3875a949ULL, 0x3fef4f87080d89f2ULL,
0x3c74a385a63d07a7ULL, 0x3fef5818dcfba487ULL,
0xbc82919e2040220fULL, 0x3fef60e316c98398ULL,
0x3c8e5a50d5c192acULL, 0x3fef69e603db3285ULL,
0x3c843a59ac016b4bULL, 0x3fef7321f301b460ULL,
0xbc82d52107b43e1fULL, 0x3fef7c97337b9b5fULL,
0xbc892ab93b470dc9ULL, 0x3fef864614f5a129ULL,
0x3c74b604603a88d3ULL, 0x3fef902ee78b3ff6ULL,
0x3c83c5ec519d7271ULL, 0x3fef9a51fbc74c83ULL,
0xbc8ff7128fd391f0ULL, 0x3fefa4afa2a490daULL,
0xbc8dae98e223747dULL, 0x3fefaf482d8e67f1ULL,
0x3c8ec3bc41aa2008ULL, 0x3fefba1bee615a27ULL,
0x3c842b94c3a9eb32ULL, 0x3fefc52b376bba97ULL,
0x3c8a64a931d185eeULL, 0x3fefd0765b6e4540ULL,
0xbc8e37bae43be3edULL, 0x3fefdbfdad9cbe14ULL,
0x3c77893b4d91cd9dULL, 0x3fefe7c1819e90d8ULL,
0x3c5305c14160cc89ULL, 0x3feff3c22b8f71f1ULL
};
static const uint64_t kExpTabHost[256] = {
0x0000000000000000ULL, 0x3ff0000000000000ULL,
0x3c9b3b4f1a88bf6eULL, 0x3feff63da9fb3335ULL,
0xbc7160139cd8dc5dULL, 0x3fefec9a3e778061ULL,
0xbc905e7a108766d1ULL, 0x3fefe315e86e7f85ULL,
0x3c8cd2523567f613ULL, 0x3fefd9b0d3158574ULL,
0xbc8bce8023f98efaULL, 0x3fefd06b29ddf6deULL,
0x3c60f74e61e6c861ULL, 0x3fefc74518759bc8ULL,
0x3c90a3e45b33d399ULL, 0x3fefbe3ecac6f383ULL,
0x3c979aa65d837b6dULL, 0x3fefb5586cf9890fULL,
0x3c8eb51a92fdeffcULL, 0x3fefac922b7247f7ULL,
0x3c3ebe3d702f9cd1ULL, 0x3fefa3ec32d3d1a2ULL,
0xbc6a033489906e0bULL, 0x3fef9b66affed31bULL,
0xbc9556522a2fbd0eULL, 0x3fef9301d0125b51ULL,
0xbc5080ef8c4eea55ULL, 0x3fef8abdc06c31ccULL,
0xbc91c923b9d5f416ULL, 0x3fef829aaea92de0ULL,
0x3c80d3e3e95c55afULL, 0x3fef7a98c8a58e51ULL,
0xbc801b15eaa59348ULL, 0x3fef72b83c7d517bULL,
0xbc8f1ff055de323dULL, 0x3fef6af9388c8deaULL,
0x3c8b898c3f1353bfULL, 0x3fef635beb6fcb75ULL,
0xbc96d99c7611eb26ULL, 0x3fef5be084045cd4ULL,
0x3c9aecf73e3a2f60ULL, 0x3fef54873168b9aaULL,
0xbc8fe782cb86389dULL, 0x3fef4d5022fcd91dULL,
0x3c8a6f4144a6c38dULL, 0x3fef463b88628cd6ULL,
0x3c807a05b0e4047dULL, 0x3fef3f49917ddc96ULL,
0x3c968efde3a8a894ULL, 0x3fef387a6e756238ULL,
0x3c875e18f274487dULL, 0x3fef31ce4fb2a63fULL,
0x3c80472b981fe7f2ULL, 0x3fef2b4565e27cddULL,
0xbc96b87b3f71085eULL, 0x3fef24dfe1f56381ULL,
0x3c82f7e16d09ab31ULL, 0x3fef1e9df51fdee1ULL,
0xbc3d219b1a6fbffaULL, 0x3fef187fd0dad990ULL,
0x3c8b3782720c0ab4ULL, 0x3fef1285a6e4030bULL,
0x3c6e149289cecb8fULL, 0x3fef0cafa93e2f56ULL,
0x3c834d754db0abb6ULL, 0x3fef06fe0a31b715ULL,
0x3c864201e2ac744cULL, 0x3fef0170fc4cd831ULL,
0x3c8fdd395dd3f84aULL, 0x3feefc08b26416ffULL,
0xbc86a3803b8e5b04ULL, 0x3feef6c55f929ff1ULL,
0xbc924aedcc4b5068ULL, 0x3feef1a7373aa9cbULL,
0xbc9907f81b512d8eULL, 0x3feeecae6d05d866ULL,
0xbc71d1e83e9436d2ULL, 0x3feee7db34e59ff7ULL,
0xbc991919b3ce1b15ULL, 0x3feee32dc313a8e5ULL,
0x3c859f48a72a4c6dULL, 0x3feedea64c123422ULL,
0xbc9312607a28698aULL, 0x3feeda4504ac801cULL,
0xbc58a78f4817895bULL, 0x3feed60a21f72e2aULL,
0xbc7c2c9b67499a1bULL, 0x3feed1f5d950a897ULL,
0x3c4363ed60c2ac11ULL, 0x3feece086061892dULL,
0x3c9666093b0664efULL, 0x3feeca41ed1d0057ULL,
0x3c6ecce1daa10379ULL, 0x3feec6a2b5c13cd0ULL,
0x3c93ff8e3f0f1230ULL, 0x3feec32af0d7d3deULL,
0x3c7690cebb7aafb0ULL, 0x3feebfdad5362a27ULL,
0x3c931dbdeb54e077ULL, 0x3feebcb299fddd0dULL,
0xbc8f94340071a38eULL, 0x3feeb9b2769d2ca7ULL,
0xbc87deccdc93a349ULL, 0x3feeb6daa2cf6642ULL,
0xbc78dec6bd0f385fULL, 0x3feeb42b569d4f82ULL,
0xbc861246ec7b5cf6ULL, 0x3feeb1a4ca5d920fULL,
0x3c93350518fdd78eULL, 0x3feeaf4736b527daULL,
0x3c7b98b72f8a9b05ULL, 0x3feead12d497c7fdULL,
0x3c9063e1e21c5409ULL, 0x3feeab07dd485429ULL,
0x3c34c7855019c6eaULL, 0x3feea9268a5946b7ULL,
0x3c9432e62b64c035ULL, 0x3feea76f15ad2148ULL,
0xbc8ce44a6199769fULL, 0x3feea5e1b976dc09ULL,
0xbc8c33c53bef4da8ULL, 0x3feea47eb03a5585ULL,
0xbc845378892be9aeULL, 0x3feea34634ccc320ULL,
0xbc93cedd78565858ULL, 0x3feea23882552225ULL,
0x3c5710aa807e1964ULL, 0x3feea155d44ca973ULL,
0xbc93b3efbf5e2228ULL, 0x3feea09e667f3bcdULL,
0xbc6a12ad8734b982ULL, 0x3feea012750bdabfULL,
0xbc6367efb86da9eeULL, 0x3fee9fb23c651a2fULL,
0xbc80dc3d54e08851ULL, 0x3fee9f7df9519484ULL,
0xbc781f647e5a3ecfULL, 0x3fee9f75e8ec5f74ULL,
0xbc86ee4ac08b7db0ULL, 0x3fee9f9a48a58174ULL,
0xbc8619321e55e68aULL, 0x3fee9feb564267c9ULL,
0x3c909ccb5e09d4d3ULL, 0x3feea0694fde5d3fULL,
0xbc7b32dcb94da51dULL, 0x3feea11473eb0187ULL,
0x3c94ecfd5467c06bULL, 0x3feea1ed0130c132ULL,
0x3c65ebe1abd66c55ULL, 0x3feea2f336cf4e62ULL,
0xbc88a1c52fb3cf42ULL, 0x3feea427543e1a12ULL,
0xbc9369b6f13b3734ULL, 0x3feea589994cce13ULL,
0xbc805e843a19ff1eULL, 0x3feea71a4623c7adULL,
0xbc94d450d872576eULL, 0x3feea8d99b4492edULL,
0x3c90ad675b0e8a00ULL, 0x3feeaac7d98a6699ULL,
0x3c8db72fc1f0eab4ULL, 0x3feeace5422aa0dbULL,
0xbc65b6609cc5e7ffULL, 0x3feeaf3216b5448cULL,
0x3c7bf68359f35f44ULL, 0x3feeb1ae99157736ULL,
0xbc93091fa71e3d83ULL, 0x3feeb45b0b91ffc6ULL,
0xbc5da9b88b6c1e29ULL, 0x3feeb737b0cdc5e5ULL,
0xbc6c23f97c90b959ULL, 0x3feeba44cbc8520fULL,
0xbc92434322f4f9aaULL, 0x3feebd829fde4e50ULL,
0xbc85ca6cd7668e4bULL, 0x3feec0f170ca07baULL,
0x3c71affc2b91ce27ULL, 0x3feec49182a3f090ULL,
0x3c6dd235e10a73bbULL, 0x3feec86319e32323ULL,
0xbc87c50422622263ULL, 0x3feecc667b5de565ULL,
0x3c8b1c86e3e231d5ULL, 0x3feed09bec4a2d33ULL,
0xbc91bbd1d3bcbb15ULL, 0x3feed503b23e255dULL,
0x3c90cc319cee31d2ULL, 0x3feed99e1330b358ULL,
0x3c8469846e735ab3ULL, 0x3feede6b5579fdbfULL,
0xbc82dfcd978e9db4ULL, 0x3feee36bbfd3f37aULL,
0x3c8c1a7792cb3387ULL, 0x3feee89f995ad3adULL,
0xbc907b8f4ad1d9faULL, 0x3feeee07298db666ULL,
0xbc55c3d956dcaebaULL, 0x3feef3a2b84f15fbULL,
0xbc90a40e3da6f640ULL, 0x3feef9728de5593aULL,
0xbc68d6f438ad9334ULL, 0x3feeff76f2fb5e47ULL,
0xbc91eee26b588a35ULL, 0x3fef05b030a1064aULL,
0x3c74ffd70a5fddcdULL, 0x3fef0c1e904bc1d2ULL,
0xbc91bdfbfa9298acULL, 0x3fef12c25bd71e09ULL,
0x3c736eae30af0cb3ULL, 0x3fef199bdd85529cULL,
0x3c8ee3325c9ffd94ULL, 0x3fef20ab5fffd07aULL,
0x3c84e08fd10959acULL, 0x3fef27f12e57d14bULL,
0x3c63cdaf384e1a67ULL, 0x3fef2f6d9406e7b5ULL,
0x3c676b2c6c921968ULL, 0x3fef3720dcef9069ULL,
0xbc808a1883ccb5d2ULL, 0x3fef3f0b555dc3faULL,
0xbc8fad5d3ffffa6fULL, 0x3fef472d4a07897cULL,
0xbc900dae3875a949ULL, 0x3fef4f87080d89f2ULL,
0x3c74a385a63d07a7ULL, 0x3fef5818dcfba487ULL,
0xbc82919e2040220fULL, 0x3fef60e316c98398ULL,
0x3c8e5a50d5c192acULL, 0x3fef69e603db3285ULL,
0x3c843a59ac016b4bULL, 0x3fef7321f301b460ULL,
0xbc82d52107b43e1fULL, 0x3fef7c97337b9b5fULL,
0xbc892ab93b470dc9ULL, 0x3fef864614f5a129ULL,
0x3c74b604603a88d3ULL, 0x3fef902ee78b3ff6ULL,
0x3c83c5ec519d7271ULL, 0x3fef9a51fbc74c83ULL,
0xbc8ff7128fd391f0ULL, 0x3fefa4afa2a490daULL,
0xbc8dae98e223747dULL, 0x3fefaf482d8e67f1ULL,
0x3c8ec3bc41aa2008ULL, 0x3fefba1bee615a27ULL,
0x3c842b94c3a9eb32ULL, 0x3fefc52b376bba97ULL,
0x3c8a64a931d185eeULL, 0x3fefd0765b6e4540ULL,
0xbc8e37bae43be3edULL, 0x3fefdbfdad9cbe14ULL,
0x3c77893b4d91cd9dULL, 0x3fefe7c1819e90d8ULL,
0x3c5305c14160cc89ULL, 0x3feff3c22b8f71f1ULL
};

BC_HD double as_double(uint64_t u) { return __builtin_bit_cast(double, u); }
BC_HD uint64_t as_u64(double d) { return __builtin_bit_cast(uint64_t, d); }

// On the device the table is read from LDS: every kernel that evaluates exp stages it once
// per workgroup with exp_table_stage() before its first exp.
__device__ static inline uint64_t* exp_table_lds() {
    __shared__ uint64_t t[256];
    return t;
}
__device__ static inline void exp_table_stage() {
    uint64_t* t = exp_table_lds();
    for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x) t[i] = kExpTabDev[i];
    __syncthreads();
}

BC_HD uint64_t exp_tab(uint32_t i) {
#if defined(__HIP_DEVICE_COMPILE__)
    return exp_table_lds()[i];
#else
    return kExpTabHost[i];
#endif
}

// Table sources for exp_t: the staged LDS copy (host: the host table), or the global copy for
// kernels that keep their LDS for other data.
struct ExpTabLds {
    __host__ __device__ uint64_t operator()(uint32_t i) const { return exp_tab(i); }
};
struct ExpTabGlobal {
#if defined(__HIP_DEVICE_COMPILE__)
    __device__ uint64_t operator()(uint32_t i) const { return kExpTabDev[i]; }
#else
    uint64_t operator()(uint32_t i) const { return kExpTabHost[i]; }
#endif
};

// Large-|x| tail of exp (|x| >= 512): rescale so the exponent cannot overflow or lose
// subnormal precision before the final multiply.
BC_HD double exp_special(double tmp, uint64_t sbits, uint64_t ki) {
    if ((ki & 0x80000000u) == 0) {
        sbits -= 1009ull << 52;
        double scale = as_double(sbits);
        return 0x1p1009 * __builtin_fma(scale, tmp, scale);
    }
    sbits += 1022ull << 52;
    double scale = as_double(sbits);
    double st = scale * tmp;
    double y = scale + st;
    if (y < 1.0) {
        double lo = scale - y + scale * tmp;
        double hi = 1.0 + y;
        lo = 1.0 - hi + y + lo;
        y = (hi + lo) - 1.0;
        if (y == 0.0) y = 0.0;
    }
    return 0x1p-1022 * y;
}

template <class Tab>
BC_HD double exp_t(double x, Tab tab) {
    const double kInvLn2N = 0x1.71547652b82fep0 * 128.0;
    const double kNegLn2HiN = -0x1.62e42fefa0000p-8;
    const double kNegLn2LoN = -0x1.cf79abc9e3b3ap-47;
    const double kShift = 0x1.8p52;
    const double C2 = 0x1.ffffffffffdbdp-2, C3 = 0x1.555555555543cp-3;
    const double C4 = 0x1.55555cf172b91p-5, C5 = 0x1.1111167a4d017p-7;
    uint32_t abstop = (uint32_t)(as_u64(x) >> 52) & 0x7ff;
    // 0x3c9 = top12(2^-54), 0x408 = top12(512), 0x409 = top12(1024)
    if (abstop - 0x3c9u >= 0x408u - 0x3c9u) {
        if ((int32_t)(abstop - 0x3c9u) < 0) return 1.0 + x;
        if (abstop >= 0x409u) {
            if (as_u64(x) == 0xfff0000000000000ull) return 0.0;
            if (abstop >= 0x7ffu) return 1.0 + x;
            return (as_u64(x) >> 63) ? 0.0 : __builtin_inf();
        }
        abstop = 0;
    }
    double kd = __builtin_fma(kInvLn2N, x, kShift);
    uint64_t ki = as_u64(kd);
    kd -= kShift;
    double r = __builtin_fma(kd, kNegLn2LoN, __builtin_fma(kd, kNegLn2HiN, x));
    uint32_t idx = 2u * (uint32_t)(ki % 128u);
    uint64_t top = ki << 45;
    double tail = as_double(tab(idx));
    uint64_t sbits = tab(idx + 1) + top;
    double r2 = r * r;
    double tmp = __builtin_fma(r2 * r2, __builtin_fma(r, C5, C4),
                               __builtin_fma(r2, __builtin_fma(r, C3, C2), tail + r));
    if (abstop == 0) return exp_special(tmp, sbits, ki);
    double scale = as_double(sbits);
    return __builtin_fma(scale, tmp, scale);
}
BC_HD double exp(double x) { return exp_t(x, ExpTabLds{}); }

// a / b, correctly rounded, given y = rcp_prep(b) (an IEEE division computed once for a divisor
// reused many times). Two FMA Newton corrections make the quotient faithful; with y correctly
// rounded, Markstein's theorem makes the last correction return RN(a / b). Zero, subnormal,
// infinite and NaN results, subnormal or zero numerators and non-normal divisors (y = NaN) take
// the IEEE division, so no underflow/overflow assumption is made.
// tests/test_div_rcp.py checks bit identity with IEEE division on 10^8 random operands.
// The IEEE division these helpers fall back to must stay on its (rare) branch: the compiler
// would otherwise evaluate it speculatively on every call and select. An empty volatile asm on
// the dividend cannot be hoisted out of the branch.
BC_HD float keep_branch(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    __asm__ volatile("" : "+v"(x));
#endif
    return x;
}
BC_HD double keep_branch(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    __asm__ volatile("" : "+v"(x));
#endif
    return x;
}

BC_HD float rcp_prep(float b) {   // RN(1/b), or NaN when b is not a normal number
    float ab = __builtin_fabsf(b);
    return (ab >= 1.17549435e-38f && ab <= 3.40282347e+38f) ? 1.0f / b : __builtin_nanf("");
}
BC_HD float div_rcp(float a, float b, float y) {
    float q0 = a * y;
    float r0 = __builtin_fmaf(-b, q0, a);
    float q1 = __builtin_fmaf(r0, y, q0);
    float r1 = __builtin_fmaf(-b, q1, a);
    float q2 = __builtin_fmaf(r1, y, q1);
    float aq = __builtin_fabsf(q2), aa = __builtin_fabsf(a);
    if (!(aq >= 1.17549435e-38f && aq <= 3.40282347e+38f && aa >= 1.17549435e-38f)) return keep_branch(a) / b;
    return q2;
}
BC_HD double rcp_prep(double b) {   // RN(1/b), or NaN when b is not a normal number
    double ab = __builtin_fabs(b);
    return (ab >= 2.2250738585072014e-308 && ab <= 1.7976931348623157e308) ? 1.0 / b : __builtin_nan("");
}
BC_HD double div_rcp(double a, double b, double y) {
    double q0 = a * y;
    double r0 = __builtin_fma(-b, q0, a);
    double q1 = __builtin_fma(r0, y, q0);
    double r1 = __builtin_fma(-b, q1, a);
    double q2 = __builtin_fma(r1, y, q1);
    double aq = __builtin_fabs(q2), aa = __builtin_fabs(a);
    if (!(aq >= 2.2250738585072014e-308 && aq <= 1.7976931348623157e308 && aa >= 2.2250738585072014e-308))
        return keep_branch(a) / b;
    return q2;
}

// a / b for floats, correctly rounded, given rb = RN(1 / (double)b). The product a * rb in
// double is within 2^-52 relative of q = a / b. A quotient of two 24-bit significands is at
// least 2^-49 (relative) away from every float rounding midpoint of the normal range (the
// overflow threshold included) and never on one, and at least 2^-48 away from the midpoint
// just below the smallest normal unless exactly on it; so whenever the product rounds to a
// normal float, that float is RN(q). Every other rounded product (zero, subnormal, infinite,
// NaN) takes the IEEE division.
BC_HD float div_wide(float a, float b, double rb) {
    float r = (float)((double)a * rb);
#if defined(__HIP_DEVICE_COMPILE__)
    if (!__builtin_amdgcn_classf(r, 0x108)) r = keep_branch(a) / b;   // class mask: -normal | +normal
#else
    if (!__builtin_isnormal(r)) r = a / b;
#endif
    return r;
}

// a / b in double, correctly rounded, for a divisor b whose significand has at most 24 bits (a
// widened float) and rb = RN(1 / b). Let a = A 2^x (A a 53-bit integer), b = B 2^y (B < 2^24)
// and M a rounding midpoint (an odd 54-bit significand). a / b = M would need A 2^k =
// (2j+1) B with 2j+1 >= 2^53 > odd(A): impossible; and a - M b is a nonzero multiple of the
// lattice unit, so a / b lies at least 2^-76 (relative) from every midpoint. q0 = RN(a rb) is
// within ~2^-52 of a / b and the corrected q0 + (a - b q0) rb within ~2^-104 as long as the
// remainder a - b q0 (~2^-53 |a|) is not subnormal, so both round alike: the result is
// RN(a / b) whenever it is a normal number and |a| >= 2^-968. Other cases (zero, subnormal,
// infinite or NaN results, tiny numerators) take the IEEE division, which also fixes the sign
// of a zero quotient.
BC_HD double div_f24(double a, double b, double rb) {
    const double q0 = a * rb;
    const double q = __builtin_fma(__builtin_fma(-b, q0, a), rb, q0);
#if defined(__HIP_DEVICE_COMPILE__)
    if (!(__builtin_amdgcn_class(q, 0x108) && __builtin_fabs(a) >= 0x1p-968)) return keep_branch(a) / b;
#else
    if (!(__builtin_isnormal(q) && __builtin_fabs(a) >= 0x1p-968)) return a / b;
#endif
    return q;
}

}  // namespace bc
