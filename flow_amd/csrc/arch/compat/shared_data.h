/* shared_data.h -- arch's shared_data.h as seen by flow's sources (included as "../shared_data.h" from
 * the flow directory): every declaration the hip kernel set provides in its
 * place is in flow_arch.h. */
#include "flow_arch.h"
