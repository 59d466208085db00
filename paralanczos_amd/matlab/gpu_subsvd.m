function [eigs, det1, err1] = gpu_subsvd(g, subIndex)
%GPU_SUBSVD  Eigenvalues of one principal submatrix of g on the GPU.
%   [eigs, det1, err1] = gpu_subsvd(g, subIndex) returns the ascending
%   eigenvalues of g(subIndex+1, subIndex+1) (subIndex is 0-based, of dyadic
%   length <= 128), their product and an error flag (0 = ok). On an invalid
%   argument it prints the reason and returns -1, -1, -1.
%   Same signature, checks and outputs as Matlab/gpu_subsvd.m:14-59.
eigs = -1; det1 = -1; err1 = -1;
[M, N] = size(g);
L = length(subIndex);
if M ~= N
    disp('dimensions of input for matrix don''t agree! ');
    return;
end
if L < 1 || floor(log2(L)) - log2(L) ~= 0
    disp('length of index is not dyadic!');
    return;
end
if L > 128
    disp('subindex of length > 128 not supported! ');
    return;
end
[eigs, det1, err1] = pl_multi_subsvd_call(g, subIndex, L);
end
