/* Export macro for the TAPP C ABI.
 * Same meaning as api/include/tapp/util.h:4-9 of the reference: default ELF
 * visibility, and extern "C" linkage when included from C++. */
#ifndef TAPP_UTIL_H_
#define TAPP_UTIL_H_

#ifdef __cplusplus
#define TAPP_EXPORT extern "C" __attribute__((visibility("default")))
#else
#define TAPP_EXPORT __attribute__((visibility("default")))
#endif

#endif /* TAPP_UTIL_H_ */
