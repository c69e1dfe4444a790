// Status codes shared by the C ABI and the host code (values match
// include/gsdr.h).
#pragma once

#define GSDR_OK 0
#define GSDR_E_INVALID 1     // bad argument or option
#define GSDR_E_IO 2          // file could not be read or written
#define GSDR_E_INPUT 3       // malformed or unsupported input
#define GSDR_E_DEVICE 4      // HIP runtime error
#define GSDR_E_NOMEM 5       // device or host allocation failed
#define GSDR_E_LIMIT 6       // a size limit of the implementation was exceeded
#define GSDR_E_BUSY 7        // serving: no slot or memory available now
#define GSDR_E_NOTFOUND 8    // serving: unknown job id
