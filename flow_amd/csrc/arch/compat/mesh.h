/* mesh.h -- arch's mesh.h as seen by flow's sources (included as "../mesh.h" from
 * the flow directory): every declaration the hip kernel set provides in its
 * place is in flow_arch.h. */
#include "flow_arch.h"
