function [eigs, det1, err1] = gpu_multi_subsvd(g, subIndex, minor)
%GPU_MULTI_SUBSVD  Eigenvalues of many principal submatrices of g on the GPU.
%   [eigs, det1, err1] = gpu_multi_subsvd(g, subIndex, minor): subIndex holds
%   0-based indices, minor (dyadic, <= 128) at a time forming
%   length(subIndex)/minor subsets. eigs is length(subIndex) x 1 (ascending
%   per subset), det1 and err1 are (length(subIndex)/minor) x 1. On an
%   invalid argument it prints the reason and returns -1, -1, -1.
%   Same signature, checks and outputs as Matlab/gpu_multi_subsvd.m:13-68.
eigs = -1; det1 = -1; err1 = -1;
[M, N] = size(g);
L = length(subIndex);
if M ~= N
    disp('dimensions of input for matrix don''t agree! ');
    return;
end
if minor < 1 || floor(log2(minor)) - log2(minor) ~= 0
    disp('length of minor is not dyadic!');
    return;
end
if mod(L, minor) ~= 0
    disp('minor must divide into the length of the index! ');
    return;
end
if minor > 128
    disp('minor of length > 128 not supported! ');
    return;
end
[eigs, det1, err1] = pl_multi_subsvd_call(g, subIndex, minor);
end
